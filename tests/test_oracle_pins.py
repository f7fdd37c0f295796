"""Pins of the FP64 CPU oracle and of the host factorisation against what the
reference publishes or what mathematics fixes.  CPU only."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from pd_ipc_amd import Cloth, Engine, SimConfig, scenes
from oracle_ffi import Oracle


def _perm(cloth, cfg):
    e = Engine("fp64")
    e.add(cloth, cfg)
    return e, e.factor_arrays(0)


def test_bench_scene_published_physics():
    """pd_bench realtime, 40x40 / kappa 1e4 / 300 substeps (docs/perf.md:1423,
    :888): final residual, final max/mean strain, run-max strain, elastic and
    total energy, 0/300 converged, 6 colours."""
    c, cfg = scenes.bench_scene_40()
    _, f = _perm(c, cfg)
    o = Oracle(c, cfg, f["perm3"], threads=4)
    worst = 0.0
    for _ in range(300):
        o.step(1)
        worst = max(worst, o.max_strain())
    assert f"{o.last_residual():.6g}" == "8.63991"
    assert f"{o.max_strain():.6g}" == "0.0585735"
    assert f"{o.mean_strain():.6g}" == "0.0013502"
    assert f"{worst:.6g}" == "0.0652267"
    assert f"{o.surrogate_energy():.6g}" == "0.322049"
    assert f"{o.total_energy():.6g}" == "-10.5941"
    s = o.stats()
    assert s["early_exit_count"] == 0 and s["factorize_count"] == 1 and s["color_count"] == 6
    assert s["compact_ready"] == 1 and s["compact_nnz"] == 22009


# pd_bench 40x40 / kappa 1e4 / 300 substeps variants as published, each figure at
# the digits it was printed with: (iterations, residual tol, bend k) ->
# residual, final max strain, mean strain, elastic energy
PUBLISHED_VARIANTS = [
    # presets preview / accurate (docs/perf.md:1422, :1424)
    ((20, 0.3, 0.0), ("13.3816", "0.0581053", None, "0.320141"), 6),
    ((80, 1e-3, 0.0), ("6.40876", "0.0586015", None, "0.322604"), 6),
    # iteration sweep (docs/solver-feasibility.md:141-144): residual printed as 24.10
    ((10, 1e-3, 0.0), ("24.10", "0.057035", "0.00134129", "0.318593"), 6),
    ((20, 1e-3, 0.0), ("13.38", "0.0581053", "0.00134602", "0.320141"), 6),
    ((80, 1e-3, 0.0), ("6.41", "0.0586015", "0.00135108", "0.322604"), 6),
    # bend stiffness sweep, PD_BEND=k (docs/perf.md:856-862)
    ((40, 1e-3, 0.001), ("8.62434", "0.0585737", "0.0013502", "0.322056"), 6),
    ((40, 1e-3, 1.0), ("50.33", "0.05872", "0.00135012", "0.326154"), 6),
    ((40, 1e-3, 100.0), ("2153.4", "0.06213", "0.00143839", "0.37633"), 6),
    ((40, 1e-3, 1000.0), ("5235.6", "0.04993", "0.00139981", "0.42291"), 6),
]


def _printed(value, text):
    """value printed with as many significant digits (or decimals) as text has."""
    if "." in text and text.endswith("0"):
        return f"{value:.{len(text.split('.')[1])}f}" == text
    digits = len(text.replace("-", "").replace(".", "").lstrip("0"))
    return f"{value:.{digits}g}" == text


@pytest.mark.parametrize("setup,expect,threads", PUBLISHED_VARIANTS,
                         ids=[f"it{s[0]}_tol{s[1]}_bend{s[2]}" for s, _, _ in PUBLISHED_VARIANTS])
def test_bench_scene_published_variants(setup, expect, threads):
    iters, rtol, bend = setup
    c = scenes.grid_cloth(40, 40, 1.0e4, bend=bend, sx=0.025, sz=0.025, mass="edges")
    cfg = SimConfig(dt=1.0 / 120.0, max_iterations=iters, residual_tol=rtol)
    _, f = _perm(c, cfg)
    o = Oracle(c, cfg, f["perm3"], threads=threads)
    o.step(300)
    res, smax, smean, el = expect
    assert _printed(o.last_residual(), res), o.last_residual()
    assert _printed(o.max_strain(), smax), o.max_strain()
    if smean is not None:
        assert _printed(o.mean_strain(), smean), o.mean_strain()
    assert _printed(o.surrogate_energy(), el), o.surrogate_energy()
    assert o.stats()["early_exit_count"] == 0  # 0/300 converged in every published run


def test_bench_200_published_realtime():
    """pd_bench --grid 200 200 --steps 100 --preset realtime (docs/perf.md:1431)."""
    c = scenes.grid_cloth(200, 200, 1.0e4, sx=0.025, sz=0.025, mass="edges")
    cfg = SimConfig(dt=1.0 / 120.0, max_iterations=40, residual_tol=0.3)
    _, f = _perm(c, cfg)
    o = Oracle(c, cfg, f["perm3"], threads=8)
    o.step(100)
    assert f"{o.last_residual():.6g}" == "58.3972"
    assert f"{o.max_strain():.6g}" == "0.529022"
    assert f"{o.surrogate_energy():.5g}" == "38.989"


def test_printed_digit_helper():
    assert _printed(24.0984, "24.10") and not _printed(24.0884, "24.10")
    assert _printed(5235.55, "5235.6") and _printed(0.0499347, "0.04993")
    assert not _printed(0.0499347, "0.04994")


@pytest.mark.parametrize("n,bend,nnz3,levels", [(40, 0.0, 66027, 155), (100, 0.0, 637233, 521),
                                                (200, 0.0, 3255033, 1125), (40, 1000.0, 182613, None)])
def test_factor_nnz_and_levels_match_published(n, bend, nnz3, levels):
    """AMD on the 3n pattern + LDL^T: factor nnz of docs/perf.md:575-577 and :980,
    sweep levels of docs/solver-feasibility.md:63-65."""
    c = scenes.grid_cloth(n, n, 1.0e4, bend=bend, sx=0.025, sz=0.025, mass="edges")
    e = Engine("fp64")
    e.add(c, SimConfig())
    st = e.stats(0) if False else None
    f = e.factor_arrays(0)
    assert 3 * len(f["rows"]) == nnz3
    s = e.stats(0)
    if levels is not None:
        assert s["fwd_levels"] == levels and s["bwd_levels"] == levels


@pytest.mark.parametrize("n,colors", [(2, 2), (5, 5), (20, 6), (40, 6)])
def test_colour_counts_published(n, colors):
    """README.md:154: 2x2 -> 2, 5x5 -> 5, 20x20 and 40x40 -> 6 colours."""
    c = scenes.grid_cloth(n, n, 1.0e4, sx=0.025, sz=0.025, mass="edges")
    _, f = _perm(c, SimConfig())
    o = Oracle(c, SimConfig(), f["perm3"])
    col, order, start = o.coloring()
    assert len(start) - 1 == colors
    # proper: no two edges of one colour share a vertex
    for k in range(colors):
        es = order[start[k]:start[k + 1]]
        verts = np.concatenate([c.edge_a[es], c.edge_b[es]])
        assert len(np.unique(verts)) == len(verts)


def _spring(y0, v0, mass, rest, k, h, iters=60):
    pos = np.array([0.0, 0.0, 0.0, 0.0, y0, 0.0])
    cl = Cloth(positions=pos, masses=[mass, mass], pinned=[1, 0], edge_a=[0], edge_b=[1],
               edge_rest=[rest], edge_kappa=[k], velocities=[0, 0, 0, 0, v0, 0])
    cfg = SimConfig(dt=h, damping=0.0, max_iterations=iters, rel_tol=0.0, abs_tol=1e-16, residual_tol=0.0)
    return cl, cfg


@pytest.mark.parametrize("k,h,y0", [(1.0, 1 / 60, 0.12), (100.0, 1 / 60, 0.12), (1e4, 1 / 120, 0.12),
                                    (1e4, 1 / 240, 0.12), (1e4, 1 / 120, 0.09), (100.0, 1 / 120, 0.08)])
def test_two_vertex_spring_closed_form(k, h, y0):
    """tests/chain/spring_vertical.cpp:60-63, :101-125: one step lands on
    y = (m yhat/h^2 + k l)/(m/h^2 + k) (y0 > l: stretched, y0 < l: compressed)."""
    mass, rest = 0.05, 0.1
    cl, cfg = _spring(-y0, 0.0, mass, rest, k, h)
    o = Oracle(cl, cfg, np.arange(6, dtype=np.int32))
    o.step(1)
    yhat = -y0 - h * h * 9.81
    mh = mass / (h * h)
    # the spring hangs below the pin: y < 0, length |y|
    expected = -(mh * (-yhat) + k * rest) / (mh + k)
    assert abs(o.positions()[4] - expected) / rest < 1e-12


def test_two_vertex_spring_equilibrium():
    """spring_vertical.cpp:162-200: long run settles at l + m g / k below the pin."""
    mass, rest, k, h = 0.05, 0.1, 1e4, 1 / 120
    cl, cfg = _spring(-0.1, 0.0, mass, rest, k, h)
    cfg.damping = 0.02
    o = Oracle(cl, cfg, np.arange(6, dtype=np.int32))
    o.step(3000)
    assert abs(o.positions()[4] - (-(rest + mass * 9.81 / k))) < 1e-9


def test_free_fall():
    """_verify/check.cpp:119-138: an unconstrained vertex moves by v h - h^2 g per step."""
    h = 1 / 120
    cl = Cloth(positions=np.zeros(3), masses=[1.0], pinned=[0], edge_a=[], edge_b=[], edge_rest=[],
               edge_kappa=[])
    cfg = SimConfig(dt=h, damping=0.0, max_iterations=5, rel_tol=0.0, abs_tol=0.0, residual_tol=0.0)
    o = Oracle(cl, cfg, np.arange(3, dtype=np.int32))
    o.step(1)
    assert abs(o.positions()[1] - (-h * h * 9.81)) < 1e-15
    o.step(1)
    assert abs(o.positions()[1] - (-3 * h * h * 9.81)) < 1e-15


@pytest.mark.parametrize("nx,ny,shear,bend", [(12, 9, 0.0, 0.0), (15, 11, 500.0, 50.0)])
def test_solve_against_scipy(nx, ny, shear, bend):
    """EigenDirectSolver::solve residual pin (spring_vertical.cpp:388-426) and an
    independent sparse LU solve of the assembled system."""
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    cfg = SimConfig(dt=1 / 600)
    _, f = _perm(c, cfg)
    o = Oracle(c, cfg, f["perm3"])
    cp, r, v = o.matrix()
    A = sp.csc_matrix((v, r, cp), shape=(c.n, c.n))
    assert abs(A - A.T).max() == 0.0
    L = sp.kron(A, sp.identity(3), format="csc")
    b = np.random.default_rng(3).standard_normal(3 * c.n)
    x = o.solve(b)
    assert np.abs(L @ x - b).max() / max(1.0, np.abs(b).max()) < 1e-12
    xs = spla.spsolve(L, b)
    assert np.abs(x - xs).max() / np.abs(xs).max() < 1e-12


def test_assembled_blocks():
    """Assembler.cpp:139-278 block rules: diagonal m/h^2 + sum kappa (free vertex),
    -kappa coupling only between free vertices, pinned rows identity; bend blocks
    k w_p w_q (test_distance_term.cpp:223-278, :961-1012)."""
    c = scenes.grid_cloth(6, 5, 2300.0, bend=50.0)
    cfg = SimConfig(dt=1 / 600)
    _, f = _perm(c, cfg)
    o = Oracle(c, cfg, f["perm3"])
    cp, r, v = o.matrix()
    A = sp.csc_matrix((v, r, cp), shape=(c.n, c.n)).toarray()
    invh2 = 1.0 / (cfg.dt * cfg.dt)
    pinned = c.pinned.astype(bool)
    for p in np.nonzero(pinned)[0]:
        assert A[p, p] == 1.0 and np.count_nonzero(A[p]) == 1 and np.count_nonzero(A[:, p]) == 1
    ref = np.zeros_like(A)
    for i in range(c.n):
        ref[i, i] += c.masses[i] * invh2
    for a, b, k in zip(c.edge_a, c.edge_b, c.edge_kappa):
        if not pinned[a]:
            ref[a, a] += k
        if not pinned[b]:
            ref[b, b] += k
        if not pinned[a] and not pinned[b]:
            ref[a, b] -= k
            ref[b, a] -= k
    w = [1.0, -2.0, 1.0]
    for s in range(c.n_bends):
        vv = [c.bend_a[s], c.bend_b[s], c.bend_c[s]]
        for p in range(3):
            for q in range(3):
                if not pinned[vv[p]] and not pinned[vv[q]]:
                    ref[vv[p], vv[q]] += c.bend_k[s] * w[p] * w[q]
    free = ~pinned
    assert np.allclose(A[np.ix_(free, free)], ref[np.ix_(free, free)], rtol=1e-14, atol=0)


def test_factor_once_stamp_rules():
    """spring_vertical.cpp:312-378 and :528-563: 1 factorisation over many steps;
    kappa / h / mass change -> +1 factorisation; pin mask change -> +1 analysis and
    +1 factorisation; moving a pin or changing damping -> +0."""
    c = scenes.grid_cloth(10, 8, 2300.0)
    cfg = SimConfig.fixed_iterations(3, 1 / 600)
    e, f = _perm(c, cfg)
    o = Oracle(c, cfg, f["perm3"])
    o.step(20)
    assert o.stats()["factorize_count"] == 1 and o.stats()["analyze_count"] == 1
    q = c.pin_positions.copy()
    q[3 * (c.n - 1)] += 0.01
    o.set_pin_positions(q)
    cfg2 = SimConfig.fixed_iterations(3, 1 / 600, damping=0.1)
    o.set_config(cfg2)
    o.step(5)
    assert o.stats()["factorize_count"] == 1
    o.set_edge_kappa(c.edge_kappa * 2)
    o.step(2)
    assert o.stats()["factorize_count"] == 2
    cfg3 = SimConfig.fixed_iterations(3, 1 / 500, damping=0.1)
    o.set_config(cfg3)
    o.step(2)
    assert o.stats()["factorize_count"] == 3
    o.set_masses(c.masses * 1.5)
    o.step(2)
    assert o.stats()["factorize_count"] == 4 and o.stats()["analyze_count"] == 1
    # pin an interior vertex where it is now: the pattern loses its coupling
    # blocks, so the order is analysed again (the product's AMD of the new pattern)
    v = 3 * 10 + 4
    x = o.positions()
    pinned = c.pinned.copy()
    pinned[v] = 1
    q2 = q.copy()
    q2[3 * v:3 * v + 3] = x[3 * v:3 * v + 3]
    e.set_pinned(0, pinned)
    f2 = e.factor_arrays(0)
    assert not np.array_equal(f2["perm3"], f["perm3"])
    o.set_pinned(pinned)
    o.set_pin_positions(q2)
    o.set_ordering(f2["perm3"])
    o.step(3)
    st = o.stats()
    assert st["factorize_count"] == 5 and st["analyze_count"] == 2
    assert np.array_equal(o.positions()[3 * v:3 * v + 3], q2[3 * v:3 * v + 3])  # held exactly


def test_product_stamp_counters_host_side():
    """The library's own counters under the same rules, host side (no device):
    pdx_get_factor refreshes the factor when a stamp changed."""
    c = scenes.grid_cloth(9, 7, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(3, 1 / 600)
    e = Engine("fp64")
    e.add(c, cfg)

    def counts():
        e.factor_arrays(0)
        s = e.stats(0)
        return s["factorize_count"], s["analyze_count"]

    assert counts() == (1, 1)
    e.set_masses(0, c.masses * 2.0)
    assert counts() == (2, 1)
    e.set_masses(0, c.masses * 2.0)  # unchanged: no refactor
    assert counts() == (2, 1)
    pinned = c.pinned.copy()
    pinned[20] = 1
    e.set_pinned(0, pinned)
    assert counts() == (3, 2)
    e.set_pinned(0, pinned)  # unchanged mask: a no-op
    assert counts() == (3, 2)
    q = c.pin_positions.copy()
    q[3 * 20 + 1] -= 0.01
    e.set_pin_positions(0, q)
    e.set_config(0, SimConfig.fixed_iterations(3, 1 / 600, damping=0.3))
    assert counts() == (3, 2)
    e.set_config(0, SimConfig.fixed_iterations(3, 1 / 500, damping=0.3))
    assert counts() == (4, 2)
    e.set_edge_kappa(0, c.edge_kappa * 3.0)
    e.set_bend_k(0, c.bend_k * 3.0)
    assert counts() == (5, 2)
    assert e.pin_count(0) == int(pinned.sum())


def test_engine_rejects_mis_sized_arrays():
    """The C side copies exactly n, 3n, E or B values: the mirror refuses other sizes."""
    c = scenes.grid_cloth(6, 5, 2300.0, bend=50.0)
    e = Engine("fp64")
    e.add(c, SimConfig.fixed_iterations(3, 1 / 600))
    n = c.n
    bad = [(e.set_state, (0, np.zeros(3 * n - 1))), (e.set_state, (0, None, np.zeros(3 * n + 3))),
           (e.set_pin_positions, (0, np.zeros(n))), (e.set_edge_kappa, (0, np.zeros(c.n_edges + 1))),
           (e.set_bend_k, (0, np.zeros(c.n_bends - 1))), (e.set_masses, (0, np.zeros(n - 1))),
           (e.set_pinned, (0, np.zeros(n + 1))), (e.solve, (0, np.zeros(3))),
           (e.set_adjoint_seed, (0, np.zeros(3 * n - 3)))]
    for fn, args in bad:
        with pytest.raises(ValueError):
            fn(*args)
    with pytest.raises(ValueError):
        e.set_pin_trajectory(0, np.zeros((4, e.pin_count(0) + 1, 3)))
    with pytest.raises(IndexError):
        e.set_masses(1, c.masses)


def test_translation_invariance():
    """convergence_criterion.cpp:161-209: shifting the whole scene shifts the
    trajectory (to rounding)."""
    c = scenes.grid_cloth(12, 10, 2300.0, shear=500.0)
    cfg = SimConfig.fixed_iterations(8, 1 / 600)
    _, f = _perm(c, cfg)
    o1 = Oracle(c, cfg, f["perm3"])
    o1.step(20)
    shift = np.tile([3.0, -2.0, 5.0], c.n)
    c2 = scenes.grid_cloth(12, 10, 2300.0, shear=500.0)
    c2.positions = c2.positions + shift
    c2.pin_positions = c2.positions.copy()
    o2 = Oracle(c2, cfg, f["perm3"])
    o2.step(20)
    assert np.abs(o2.positions() - shift - o1.positions()).max() < 1e-9


def test_thread_count_independence():
    """test_distance_term.cpp:465-548: results are bitwise equal for any thread count."""
    c = scenes.grid_cloth(30, 30, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(6, 1 / 600)
    _, f = _perm(c, cfg)
    outs = []
    for t in (1, 2, 4, 7):
        o = Oracle(c, cfg, f["perm3"], threads=t)
        o.step(5)
        outs.append(o.positions())
    for x in outs[1:]:
        assert np.array_equal(x.view(np.uint64), outs[0].view(np.uint64))


def _bench60(iters, rho):
    c = scenes.grid_cloth(60, 60, 1.0e4, sx=0.025, sz=0.025, mass="edges")
    cfg = SimConfig(dt=1.0 / 120.0, max_iterations=iters, residual_tol=0.3, chebyshev_rho=rho,
                    chebyshev_start=10)
    _, f = _perm(c, cfg)
    return c, cfg, f["perm3"]


@pytest.mark.parametrize("iters,plain,cheb", [(10, "27.3033", "27.3033"), (20, "12.3328", "9.37677"),
                                              (40, "8.59041", "7.40925"), (80, "6.14024", "3.51819")])
def test_bench60_published_residuals_plain_and_chebyshev(iters, plain, cheb):
    """pd_bench --grid 60 60 --steps 300 (docs/chebyshev.md:284-287): final
    residual of the plain iteration and of the Chebyshev blend (PD_CHEB=1,
    rho 0.98, start 10) at 10/20/40/80 iterations, 4 threads."""
    for rho, want in ((0.0, plain), (0.98, cheb)):
        c, cfg, perm = _bench60(iters, rho)
        o = Oracle(c, cfg, perm, threads=4)
        o.step(300)
        assert f"{o.last_residual():.6g}" == want, (iters, rho)

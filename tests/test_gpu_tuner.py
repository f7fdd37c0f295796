"""The autotuner on single cloths (fill order, launch form, shape) and the grid
form's level sweeps: choices are timed on the device, the probes leave no trace
in the simulation, and a tuned context gives the same bits as a context set by
hand to the same choice."""
import numpy as np
import pytest

from oracle_ffi import Oracle
from pd_ipc_amd import Engine, SimConfig, scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed32"])
def test_tuner_picks_form_and_leaves_no_trace(precision):
    """Config 2's cloth (100 x 100, shear + bend, perturbed): the grid form wins
    over one workgroup by a wide margin, so the tuner must choose it; the state
    and the step statistics are untouched; a context set by hand to the tuner's
    choice steps to the same bits; FP64 stays within 1e-10 of the oracle."""
    c, cfg, _ = scenes.config2()
    _say("tuner on 100x100", precision)
    a = Engine(precision, "fast")
    a.add(c, cfg)
    x0, v0 = a.state(0)
    blk, cap, ms = a.autotune(1)
    form, blocks, order = a.launch()
    print("chosen", form, blocks, order, blk, cap, ms)
    assert form == "grid" and 16 <= blocks <= 256 and ms > 0
    st = a.stats(0)
    assert st["substeps"] == 0 and st["total_iterations"] == 0
    xa, va = a.state(0)
    assert np.array_equal(xa, x0) and np.array_equal(va, v0)
    b = Engine(precision, "fast")
    b.add(c, cfg)
    b.set_ordering(order)
    b.set_launch(form, blocks)
    b.set_schedule(blk, cap)
    nsub = 3
    a.step(nsub)
    b.step(nsub)
    xa, xb = a.state(0)[0], b.state(0)[0]
    assert np.array_equal(xa.view(np.uint64), xb.view(np.uint64))
    assert a.stats(0)["substeps"] == nsub
    if precision == "fp64":
        amd = Engine("fp64", "exact")
        amd.add(c, cfg)
        o = Oracle(c, cfg, amd.factor_arrays(0)["perm3"], threads=4)
        o.step(nsub)
        xo = o.positions()
        assert np.abs(xa - xo).max() <= 1e-10 * np.abs(xo).max()
        amd.close()
    a.close()
    b.close()


def test_grid_bits_do_not_depend_on_the_workgroup_count():
    """A level sweep's arithmetic is fixed by its schedule (tasks, runs, lanes,
    panels), not by which wavefront runs a task: 16, 64 and 256 workgroups give
    the same bits."""
    c = scenes.grid_cloth(64, 57, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(12, scenes.H600)
    xs = []
    for g in (16, 64, 256):
        _say("grid form mixed32 64x57 on", g, "workgroups")
        e = Engine("mixed32", "fast")
        e.add(c, cfg)
        e.set_ordering("nd")
        e.set_launch("grid", g)
        e.step(4)
        xs.append(e.state(0)[0])
        e.close()
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])


def test_grid_solve_at_size():
    """pdx_solve3 through the grid form on the 512 x 512 cloth (long chains of up
    to 1,024 rows applied as inverse panels), FP64: the solution satisfies
    L D L^T x = b of the same factor (host sparse products) to 1e-11 relative,
    per coordinate, in rank space."""
    import scipy.sparse as sp
    c, cfg, _ = scenes.config3()
    _say("512x512 grid solve3 fp64")
    e = Engine("fp64", "fast")
    e.add(c, cfg)
    e.set_ordering("nd")
    e.set_launch("grid", 128)
    b = np.random.default_rng(5).standard_normal(3 * c.n)
    x = e.solve(0, b)
    f = e.factor_arrays(0)
    e.close()
    m = c.n
    L = sp.csc_matrix((f["vals"], f["rows"], f["ptr"]), shape=(m, m)) + sp.identity(m, format="csc")
    rv = f["rank_vertex"]
    for d in range(3):
        xr = x[3 * rv + d]
        br = b[3 * rv + d]
        back = L @ ((L.T @ xr) / f["dinv"])
        assert np.abs(back - br).max() <= 1e-11 * np.abs(br).max()

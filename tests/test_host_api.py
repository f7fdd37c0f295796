"""Host-side behaviour of the C ABI that needs no GPU: argument checks, error
codes, schedule statistics, factor-once bookkeeping.  CPU only."""
import numpy as np
import pytest

from pd_ipc_amd import Engine, PdxError, SimConfig, scenes


def _cloth():
    return scenes.grid_cloth(30, 22, 2300.0, shear=500.0, bend=50.0)


def test_schedule_stats_cover_the_factor():
    """Fast-mode schedules stream every strictly-lower factor entry once per
    sweep: runs of 16 hold at least F entries, every rank sits in a task lane."""
    c = _cloth()
    e = Engine("fp64", "fast")
    e.add(c, SimConfig.fixed_iterations(10, 1.0 / 600.0))
    f = e.factor_arrays(0)
    st = e.schedule_stats(0)
    F = len(f["rows"])
    for d in ("fwd", "bwd"):
        s = st[d]
        assert s["steps"] >= 1 and s["tasks"] >= s["steps"]
        # entries inside chain blocks are applied through the blocks' inverses,
        # the rest are streamed in runs
        assert 16 * s["runs"] <= F + 15 * c.n
        assert s["max_runs_per_step"] <= s["runs"]
        assert 0 < s["chain_rows"] <= c.n
    e.close()


def test_chain_cap_changes_schedule_not_factor():
    c = _cloth()
    cfg = SimConfig.fixed_iterations(10, 1.0 / 600.0)
    a = Engine("fp64", "fast")
    a.add(c, cfg)
    s32 = a.schedule_stats(0)
    a.set_schedule(256, 8)
    s8 = a.schedule_stats(0)
    assert s8["fwd"]["steps"] > s32["fwd"]["steps"]
    st = a.stats(0)
    assert st["factorize_count"] == 1 and st["analyze_count"] == 1  # repacked, not refactored
    a.close()


@pytest.mark.parametrize("threads,cap", [(32, 16), (1024, 16), (96, 16), (256, 0), (256, 33)])
def test_schedule_rejects_bad_shapes(threads, cap):
    e = Engine("fp64", "fast")
    e.add(_cloth(), SimConfig.fixed_iterations(10, 1.0 / 600.0))
    with pytest.raises(PdxError):
        e.set_schedule(threads, cap)
    e.close()


def test_grid_form_needs_one_instance_and_fast_mode():
    cfg = SimConfig.fixed_iterations(10, 1.0 / 600.0)
    e = Engine("fp64", "fast")
    e.add(_cloth(), cfg)
    e.add(_cloth(), cfg)
    with pytest.raises(PdxError):
        e.set_launch("grid")
    e.close()
    e = Engine("fp64", "exact")
    e.add(_cloth(), cfg)
    with pytest.raises(PdxError):
        e.set_launch("grid")
    e.close()
    e = Engine("mixed32", "fast")
    e.add(_cloth(), cfg)
    e.set_launch("grid", 32)  # accepted; runs on the GPU
    e.close()


def test_dense_mode_preconditions():
    """The dense solve mode takes one instance and no differentiable rollout;
    with one instance the host factor (and the inverse) forms without a GPU."""
    cfg = SimConfig.fixed_iterations(10, 1.0 / 600.0)
    e = Engine("fp64", "dense")
    e.add(_cloth(), cfg)
    e.add(_cloth(), cfg)
    with pytest.raises(PdxError):
        e.factor_arrays(0)
    e.close()
    e = Engine("mixed32", "dense")
    e.add(_cloth(), cfg)
    f = e.factor_arrays(0)
    assert len(f["dinv"]) == _cloth().n
    with pytest.raises(PdxError):
        e.rollout(4, 2)
    e.set_launch("grid", 32)  # only the workgroup count applies
    e.close()


def test_invalid_scenes_are_rejected():
    c = _cloth()
    c.edge_a = c.edge_a.copy()
    c.edge_a[3] = c.n + 5
    e = Engine("fp64", "fast")
    with pytest.raises(PdxError):
        e.add(c, SimConfig.fixed_iterations(10, 1.0 / 600.0))
    e.close()


def test_pin_trajectory_shape_checked():
    c = _cloth()
    e = Engine("mixed32", "fast")
    e.add(c, SimConfig.fixed_iterations(10, 1.0 / 600.0))
    assert e.pin_count(0) == 2
    with pytest.raises(ValueError):
        e.set_pin_trajectory(0, np.zeros((5, 3, 3)))
    e.set_pin_trajectory(0, np.zeros((5, 2, 3)))
    e.set_pin_trajectory(0, None)
    e.close()


@pytest.mark.gpu
def test_bytes_accounting_formula():
    """pdx_bytes_per_iteration restates the algorithmic byte model of the
    substep (factor both sweeps, maps, edges + incidence twice, vector passes,
    bend data) at storage widths."""
    c = _cloth()
    e = Engine("fp64", "exact")
    e.add(c, SimConfig.fixed_iterations(10, 1.0 / 600.0))
    f = e.factor_arrays(0)
    n, E, B, F, m = c.n, len(c.edge_a), len(c.bend_a), len(f["rows"]), c.n
    w, i = 8, 4
    exp = 2 * F * (w + i) + 2 * (m + 1) * 4 + m * (w + 4)
    exp += 2 * (E * (2 * 4 + 2 * 8) + (n + 1) * 4 + 2 * E * 4)
    exp += 24 * n * 8 + n * 8 + n
    exp += B * (3 * 4 + 8) + (n + 1) * 4 + 3 * B * (4 + 8) + 3 * n * 8
    e.factor()  # the accounting is filled at upload
    assert e.bytes_per_iteration() == pytest.approx(exp)
    assert e.bytes_per_substep() == pytest.approx(10 * exp + 27 * n * 8)
    e.close()


def test_nd_ordering_preconditions():
    cfg = SimConfig.fixed_iterations(10, 1.0 / 600.0)
    for solve in ("exact", "dense"):
        e = Engine("fp64", solve)
        e.add(_cloth(), cfg)
        with pytest.raises(PdxError):
            e.set_ordering("nd")
        e.set_ordering("amd")  # the default is always accepted
        e.close()


def test_nd_factor_reproduces_the_matrix():
    """The nested-dissection factor is a factor of the same matrix: L D L^T in rank
    space, mapped back through rank -> vertex, equals the oracle's assembled
    component matrix to rounding; its vertex order keeps the three components of a
    vertex together; it re-analyses once; it has fewer sweep steps than AMD's."""
    import scipy.sparse as sp
    from oracle_ffi import Oracle

    c = scenes.grid_cloth(60, 52, 2300.0, shear=500.0, bend=50.0)
    cfg = SimConfig.fixed_iterations(10, 1.0 / 600.0)
    e = Engine("fp64", "fast")
    e.add(c, cfg)
    amd_steps = e.schedule_stats(0)["fwd"]["steps"]
    e.set_ordering("nd")
    f = e.factor_arrays(0)
    st = e.stats(0)
    assert st["analyze_count"] == 2 and st["factorize_count"] == 2
    assert e.schedule_stats(0)["fwd"]["steps"] < amd_steps
    n = c.n
    p3 = f["perm3"]
    assert np.array_equal(np.sort(p3), np.arange(3 * n))
    assert np.array_equal(p3[0::3] // 3, f["rank_vertex"]) and np.array_equal(p3 % 3, np.tile([0, 1, 2], n))
    L = sp.csc_matrix((f["vals"], f["rows"], f["ptr"]), shape=(n, n)) + sp.identity(n)
    R = (L @ sp.diags(1.0 / f["dinv"]) @ L.T).toarray()
    cp, r, v = Oracle(c, cfg, p3).matrix()
    A = sp.csc_matrix((v, r, cp), shape=(n, n)).toarray()
    rv = f["rank_vertex"]
    assert np.abs(R - A[np.ix_(rv, rv)]).max() <= 1e-12 * np.abs(A).max()
    e.close()

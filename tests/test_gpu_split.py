"""Instances split over two workgroups (pdx_set_split) on the GPU: the two halves
of a nested-dissection split sweep on their own workgroups and meet at the
separator, so the results must be those of one workgroup bit for bit (the same
blocks, runs and sums).  Checked on fixed-iteration runs, on runs whose residual
test stops substeps early, on the stand-alone solve, and around a differentiable
rollout (which runs every instance on one workgroup).  Every GPU step prints its
case first."""
import numpy as np
import pytest

from pd_ipc_amd import Engine, SimConfig, scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def _cloths():
    out = []
    for i, (nx, ny, sh, be) in enumerate([(70, 80, 460.0, 46.0), (40, 40, 0.0, 0.0), (55, 23, 460.0, 0.0),
                                          (21, 30, 0.0, 46.0), (90, 61, 0.0, 46.0)]):
        c = scenes.grid_cloth(nx, ny, 2300.0 * (1 + 0.3 * i), shear=sh, bend=be)
        c.velocities = 0.1 * np.random.default_rng(i).standard_normal(3 * c.n)
        c.velocities.reshape(-1, 3)[c.pinned != 0] = 0.0
        out.append(c)
    return out


def _engine(cloths, cfg, split):
    e = Engine("mixed32", "fast")
    for c in cloths:
        e.add(c, cfg)
    e.set_ordering("nd")
    e.set_split(split)
    e.factor()
    return e


@pytest.mark.parametrize("residual_tol,rel_tol", [(1e-30, 0.0), (4.0, 0.0), (0.0, 2e-2)])
def test_split_equals_one_workgroup(residual_tol, rel_tol):
    """Fixed iterations (residual evaluated, never met), a residual test that
    stops substeps early (threshold 4 |g|) and the displacement test (relative
    2e-2, which needs the displacement scale of both halves): states, iteration
    counts and residuals agree bit for bit between one workgroup per instance and
    split instances."""
    cloths = _cloths()
    cfg = SimConfig(dt=scenes.H600, max_iterations=24, rel_tol=rel_tol, abs_tol=0.0, residual_tol=residual_tol)
    _say("one workgroup per instance, residual tol", residual_tol, "rel tol", rel_tol)
    a = _engine(cloths, cfg, 0)
    a.step_graph(1, 12)
    _say("split instances of >= 21000 factor entries")
    b = _engine(cloths, cfg, 21000)
    assert b.split_count() == 4
    b.step_graph(1, 12)
    early = 0
    for i in range(len(cloths)):
        xa, va = a.state(i)
        xb, vb = b.state(i)
        assert _same(xa, xb) and _same(va, vb), i
        sa, sb = a.stats(i), b.stats(i)
        for k in ("iterations_used", "converged", "early_exit_count", "total_iterations", "substeps",
                  "last_residual", "factorize_count"):
            assert sa[k] == sb[k], (i, k, sa[k], sb[k])
        early += sb["early_exit_count"]
    if residual_tol > 1e-20 or rel_tol > 0:
        assert early > 0  # the stop test did fire (both workgroups took it together)


def test_split_solve3_equals_one_workgroup():
    c = _cloths()[0]
    cfg = SimConfig.fixed_iterations(10, scenes.H600)
    b = np.random.default_rng(3).standard_normal(3 * c.n)
    _say("solve3 one workgroup")
    x1 = _engine([c], cfg, 0).solve(0, b)
    _say("solve3 split")
    e = _engine([c], cfg, 1)
    assert e.split_count() == 1
    x2 = e.solve(0, b)
    assert _same(x1, x2)


def test_split_around_a_rollout():
    """A split context runs the differentiable rollout and its adjoint on one
    workgroup per instance, then splits again for stepping: the same bits as a
    context that never splits."""
    cloths = _cloths()[:3]
    cfg = SimConfig.fixed_iterations(8, scenes.H600)
    res = []
    for split in (0, 21000):
        _say("rollout + adjoint + step, split", split)
        e = _engine(cloths, cfg, split)
        e.step(3)
        e.rollout(4, 2)
        for i, c in enumerate(cloths):
            x, _ = e.state(i)
            e.set_adjoint_seed(i, x)
        e.adjoint()
        g = [e.gradients(i)["v0"] for i in range(len(cloths))]
        e.step(3)
        res.append(([e.state(i)[0] for i in range(len(cloths))], g, e.split_count()))
        e.close()
    assert res[0][2] == 0 and res[1][2] == 3
    for xa, xb in zip(res[0][0], res[1][0]):
        assert _same(xa, xb)
    for ga, gb in zip(res[0][1], res[1][1]):
        assert _same(ga, gb)


@pytest.mark.timeout(600)
def test_split_full_mix_equals_one_workgroup():
    """The 512-instance mix, 10 substeps, with its largest cloths split: every
    instance bit for bit as on one workgroup."""
    cloths, cfg, _ = scenes.config4()
    _say("mix, one workgroup per instance")
    a = Engine("mixed32", "fast")
    b = Engine("mixed32", "fast")
    for c in cloths:
        a.add(c, cfg)
        b.add(c, cfg)
    for e, split in ((a, 0), (b, 300000)):
        e.set_ordering("nd")
        e.set_split(split)
        e.factor()
    assert b.split_count() > 8
    a.step_graph(1, 10)
    _say("mix, split", b.split_count(), "instances")
    b.step_graph(1, 10)
    print("kernel ms per launch: one workgroup", a.last_kernel_ms(), "split", b.last_kernel_ms())
    for i in range(len(cloths)):
        assert _same(a.state(i)[0], b.state(i)[0]), i


def test_autotune_chooses_split_and_leaves_no_trace():
    """The autotuner also times split thresholds; its probes are undone (no substep
    counted, the state as before) and stepping afterwards agrees with an untuned
    context to the FP32 form's rounding."""
    cloths = _cloths()
    cfg = SimConfig.fixed_iterations(12, scenes.H600)
    _say("autotune with split thresholds")
    a = _engine(cloths, cfg, 0)
    b = _engine(cloths, cfg, 0)
    blk, cap, ms = a.autotune(1)
    th = a.split_threshold()
    print("chosen", blk, cap, ms, "split threshold", th, "split instances", a.split_count())
    sizes = sorted((len(a.factor_arrays(i)["rows"]) for i in range(len(cloths))), reverse=True)
    assert th == 0 or th in sizes
    for i in range(len(cloths)):
        assert a.stats(i)["substeps"] == 0 and a.stats(i)["factorize_count"] == 1
    a.step(4)
    b.step(4)
    for i in range(len(cloths)):
        assert np.abs(a.state(i)[0] - b.state(i)[0]).max() <= 1e-6

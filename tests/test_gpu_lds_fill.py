"""No kernel reads LDS it has not written first.  A workgroup starts with whatever
the previous workgroup on its CU left in LDS; every CU's LDS is filled with a NaN
or infinity pattern (tests/liblds_fill.so) before each launch, and the results
must be those of a run without the fill, bit for bit: the substep in every
launch form, the stand-alone solve, and the differentiable rollout with its
adjoint.

The split form is the case that needed it: runs are padded with value-0
entries at column 0, rank 0 lies in the lead's half, and the partner's first
forward sweep of a launch ran before the exchange had put a value there, so a
non-finite leftover made 0 * W[0] NaN.  Every GPU step prints its case first."""
import ctypes
import os

import numpy as np
import pytest

from pd_ipc_amd import Engine, SimConfig, scenes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PATTERNS = {"nan": 0x7FC00000, "inf": 0x7F800000}


def _say(*a):
    print("[case]", *a, flush=True)


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def _fill():
    path = os.path.join(HERE, "liblds_fill.so")
    if not os.path.exists(path):
        raise RuntimeError("tests/liblds_fill.so is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    lib.lds_fill.argtypes = [ctypes.c_uint, ctypes.c_int]
    lib.lds_fill.restype = ctypes.c_int
    return lib.lds_fill


def _batch_cloths():
    out = []
    for i, (nx, ny, sh, be) in enumerate([(70, 80, 460.0, 46.0), (40, 40, 0.0, 0.0), (55, 23, 460.0, 0.0),
                                          (21, 30, 0.0, 46.0), (90, 61, 0.0, 46.0)]):
        c = scenes.grid_cloth(nx, ny, 2300.0 * (1 + 0.3 * i), shear=sh, bend=be)
        c.velocities = 0.1 * np.random.default_rng(i).standard_normal(3 * c.n)
        c.velocities.reshape(-1, 3)[c.pinned != 0] = 0.0
        out.append(c)
    return out


# (precision, solve mode, launch form, fill order, split threshold); the dense
# mode and the grid form take one cloth
FORMS = {
    "fp64-exact": ("fp64", "exact", "batch", "amd", 0),
    "fp64-fast": ("fp64", "fast", "batch", "amd", 0),
    "mixed32-fast": ("mixed32", "fast", "batch", "nd", 0),
    "mixed32-fast-split": ("mixed32", "fast", "batch", "nd", 21000),
    "fp64-dense": ("fp64", "dense", "batch", "amd", 0),
    "fp64-grid": ("fp64", "fast", "grid", "nd", 0),
    "mixed32-grid": ("mixed32", "fast", "grid", "nd", 0),
}


def _engine(form, cloths, cfg):
    prec, solve, launch, order, split = FORMS[form]
    e = Engine(prec, solve)
    for c in cloths:
        e.add(c, cfg)
    if order != "amd":
        e.set_ordering(order)
    if launch == "grid":
        e.set_launch("grid", 64)
    if split:
        e.set_split(split)
    e.factor()
    return e


@pytest.mark.parametrize("pattern", sorted(PATTERNS))
@pytest.mark.parametrize("form", list(FORMS))
def test_leftover_lds_changes_nothing(form, pattern):
    fill = _fill()
    single = FORMS[form][1] == "dense" or FORMS[form][2] == "grid"
    cloths = [_batch_cloths()[1]] if single else _batch_cloths()
    cfg = SimConfig.fixed_iterations(12, scenes.H600)
    clean = _engine(form, cloths, cfg)
    filled = _engine(form, cloths, cfg)
    if FORMS[form][4]:
        assert filled.split_count() == 4
    for k in range(3):
        _say(form, "launch", k, "without fill")
        clean.step(1)
        _say(form, "launch", k, "after filling LDS with", pattern)
        assert fill(PATTERNS[pattern], 4096) == 0
        filled.step(1)
    for i in range(len(cloths)):
        xc, vc = clean.state(i)
        xf, vf = filled.state(i)
        assert np.all(np.isfinite(xf)) and np.all(np.isfinite(vf)), (form, i)
        assert _same(xc, xf) and _same(vc, vf), (form, i, float(np.abs(xc - xf).max()))


def test_split_partner_after_lds_fill_equals_one_workgroup():
    """The split form against one workgroup per instance, with LDS filled with NaN
    before every launch: the same bits (the partner's first sweep of a launch
    meets only values it or the exchange wrote)."""
    fill = _fill()
    cloths = _batch_cloths()
    cfg = SimConfig.fixed_iterations(12, scenes.H600)
    one = _engine("mixed32-fast", cloths, cfg)
    split = _engine("mixed32-fast-split", cloths, cfg)
    assert split.split_count() == 4
    for k in range(3):
        _say("one workgroup, launch", k)
        one.step(1)
        _say("split, launch", k, "after filling LDS with NaN")
        assert fill(PATTERNS["nan"], 4096) == 0
        split.step(1)
    for i in range(len(cloths)):
        assert _same(one.state(i)[0], split.state(i)[0]), i


@pytest.mark.parametrize("form", ["fp64-exact", "fp64-fast", "mixed32-fast-split", "fp64-dense", "fp64-grid"])
def test_solve_after_lds_fill(form):
    """The stand-alone global solve (pdx_solve3) after every CU's LDS was filled
    with NaN: the bits of a solve without the fill."""
    fill = _fill()
    single = FORMS[form][1] == "dense" or FORMS[form][2] == "grid"
    cloths = [_batch_cloths()[1]] if single else _batch_cloths()
    cfg = SimConfig.fixed_iterations(12, scenes.H600)
    clean = _engine(form, cloths, cfg)
    filled = _engine(form, cloths, cfg)
    rng = np.random.default_rng(7)
    for i, c in enumerate(cloths):
        b = rng.standard_normal(3 * c.n)
        _say(form, "solve, instance", i)
        xc = clean.solve(i, b)
        assert fill(PATTERNS["nan"], 4096) == 0
        xf = filled.solve(i, b)
        assert np.all(np.isfinite(xf)), (form, i)
        assert _same(xc, xf), (form, i)


@pytest.mark.parametrize("prec,order", [("fp64", "amd"), ("mixed32", "nd")])
def test_rollout_and_adjoint_after_lds_fill(prec, order):
    """The differentiable rollout and its adjoint (rollout segments, reverse
    kernel) after every CU's LDS was filled with NaN before each call: the
    states and gradients of a run without the fill, bit for bit."""
    fill = _fill()
    cloths = _batch_cloths()[1:4]
    cfg = SimConfig.fixed_iterations(6, scenes.H600)
    nsub = 4
    out = []
    for use_fill in (False, True):
        e = Engine(prec, "fast")
        for i, c in enumerate(cloths):
            e.add(c, cfg)
        if order != "amd":
            e.set_ordering(order)
        for i, c in enumerate(cloths):
            e.set_pin_trajectory(i, scenes.pin_trajectory(c, nsub, cfg.dt))
        _say(prec, order, "rollout, fill", use_fill)
        if use_fill:
            assert fill(PATTERNS["nan"], 4096) == 0
        e.rollout(nsub, 2)
        xs = [e.state(i)[0] for i in range(len(cloths))]
        for i, x in enumerate(xs):
            e.set_adjoint_seed(i, 2.0 * x)
        _say(prec, order, "adjoint, fill", use_fill)
        if use_fill:
            assert fill(PATTERNS["nan"], 4096) == 0
        e.adjoint()
        gs = [e.gradients(i) for i in range(len(cloths))]
        e.close()
        out.append((xs, gs))
    for i in range(len(cloths)):
        assert _same(out[0][0][i], out[1][0][i]), i
        for k in out[0][1][i]:
            assert np.all(np.isfinite(out[1][1][i][k])), (i, k)
            assert _same(out[0][1][i][k], out[1][1][i][k]), (i, k)

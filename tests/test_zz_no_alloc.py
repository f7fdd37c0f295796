"""pdx_solve3_noalloc (the allocation-free solve: static device staging instead of
pdx_solve3's per-context buffers) through the C ABI, against pdx_solve3: the same
bits, on two calls in a row with the second on a smaller instance, in the exact,
fast and split forms; refused before an upload and beyond its staging size."""
import ctypes as C

import numpy as np
import pytest

from pd_ipc_amd import Engine, SimConfig, scenes
from pd_ipc_amd.pdx import load_library

pytestmark = pytest.mark.gpu

CFG = SimConfig.fixed_iterations(10, scenes.H600)


def _noalloc(e, i, b):
    fn = load_library().pdx_solve3_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    b = np.ascontiguousarray(b, np.float64)
    x = np.full_like(b, np.nan)
    rc = fn(e.h, i, b.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p))
    return rc, x


@pytest.mark.parametrize("precision,solve,split", [("fp64", "exact", 0), ("fp64", "fast", 0),
                                                   ("mixed32", "fast", 1)])
def test_noalloc_solve_matches_solve3(precision, solve, split):
    big = scenes.grid_cloth(60, 52, 2300.0, shear=460.0, bend=46.0)
    small = scenes.grid_cloth(23, 17, 900.0)
    e = Engine(precision, solve)
    e.add(big, CFG)
    e.add(small, CFG)
    if solve == "fast":
        e.set_ordering("nd")
        e.set_split(split)
    rc, _ = _noalloc(e, 0, np.zeros(3 * big.n))
    assert rc == -2  # not factored / uploaded yet: it would have to allocate
    e.factor()
    if split:
        assert e.split_count() == 2
    print("[case] noalloc", precision, solve, split, flush=True)
    for i, c, seed in ((0, big, 1), (1, small, 2)):  # the second call on the smaller instance
        b = np.random.default_rng(seed).standard_normal(3 * c.n)
        rc, x = _noalloc(e, i, b)
        assert rc == 0
        ref = e.solve(i, b)
        assert np.array_equal(x.view(np.uint64), ref.view(np.uint64)), (i, np.abs(x - ref).max())
    e.close()


def test_noalloc_solve_refuses_beyond_its_staging():
    c = scenes.grid_cloth(363, 362, 2300.0)  # 131,406 vertices > 131,072
    e = Engine("fp64", "fast")
    e.add(c, CFG)
    e.factor()
    print("[case] noalloc too large", flush=True)
    rc, _ = _noalloc(e, 0, np.zeros(3 * c.n))
    assert rc == -1
    assert b"at most 131072 vertices" in load_library().pdx_last_error(e.h)
    e.close()

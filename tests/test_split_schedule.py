"""Schedules of instances split over two workgroups, checked on the host (no
device): pdx_verify_schedule replays the packed block schedules in the kernel's
phase order - each half seeing only what its workgroup holds - and compares the
result with a direct LDL^T solve of the same factor.  CPU only."""
import pytest

from pd_ipc_amd import Engine, PdxError, SimConfig, scenes

CFG = SimConfig.fixed_iterations(10, 1.0 / 600.0)


@pytest.mark.parametrize("nx,ny,shear,bend", [(96, 94, 460.0, 46.0), (40, 33, 0.0, 0.0), (60, 17, 460.0, 0.0),
                                              (17, 90, 0.0, 46.0), (16, 16, 460.0, 46.0), (3, 41, 0.0, 46.0)])
@pytest.mark.parametrize("cap", [32, 8])
def test_split_and_single_schedules_replay(nx, ny, shear, bend, cap):
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    errs = []
    for split in (0, 1):
        e = Engine("mixed32", "fast")
        e.add(c, CFG)
        e.set_ordering("nd")
        e.set_schedule(512, cap)
        e.set_split(split)
        errs.append(e.verify_schedule(0))
        assert e.split_count() == split
        assert e.stats(0)["factorize_count"] == 1
        e.close()
    assert max(errs) <= 1e-13
    assert errs[0] == errs[1]  # the same blocks, runs and sums either way


def test_split_only_where_it_applies():
    """The split takes the FP32-solve batch form in the nested-dissection order
    for instances of at least the threshold; everything else stays on one
    workgroup (and replays as such)."""
    big = scenes.grid_cloth(50, 40, 2300.0, shear=460.0)
    small = scenes.grid_cloth(20, 20, 2300.0)
    e = Engine("mixed32", "fast")
    e.add(big, CFG)
    e.add(small, CFG)
    e.set_ordering("nd")
    e.set_split(10000)
    e.verify_schedule(0)
    e.verify_schedule(1)
    assert e.split_count() == 1
    e.set_split(0)
    e.verify_schedule(0)
    assert e.split_count() == 0
    e.close()
    for prec, order in (("fp64", "nd"), ("mixed32", "amd")):
        e = Engine(prec, "fast")
        e.add(big, CFG)
        if order == "nd":
            e.set_ordering("nd")
        e.set_split(1)
        e.verify_schedule(0)
        assert e.split_count() == 0
        e.close()
    e = Engine("mixed32", "exact")
    e.add(big, CFG)
    with pytest.raises(PdxError):
        e.verify_schedule(0)  # no block schedules in the exact mode
    e.close()

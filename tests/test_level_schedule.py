"""Level schedules of the grid form (one step per level of whole elimination-tree
chains, long chains applied as dense inverse panels), checked on the host (no
device): pdx_verify_schedule replays them in the kernel's order - phase A tasks
reading only ranks final before their step, phase B reading only R written in
its step's phase A, every index within its array - and compares the result
with a direct LDL^T solve of the same factor.  CPU only."""
import pytest

from pd_ipc_amd import Engine, SimConfig, scenes

CFG = SimConfig.fixed_iterations(10, 1.0 / 600.0)


@pytest.mark.parametrize("nx,ny,shear,bend", [(96, 94, 460.0, 46.0), (40, 40, 0.0, 0.0), (100, 100, 500.0, 50.0),
                                              (60, 17, 460.0, 0.0), (3, 41, 0.0, 46.0), (2, 2, 0.0, 0.0)])
@pytest.mark.parametrize("order", ["nd", "amd"])
@pytest.mark.parametrize("prec", ["mixed32", "fp64"])
def test_level_schedule_replays(nx, ny, shear, bend, order, prec):
    c = scenes.grid_cloth(nx, ny, 2300.0, shear=shear, bend=bend)
    e = Engine(prec, "fast")
    e.add(c, CFG)
    if order == "nd":
        e.set_ordering("nd")
    e.set_launch("grid", 32)
    err = e.verify_schedule(0)
    assert err <= 1e-12
    st = e.schedule_stats(0)
    assert e.stats(0)["factorize_count"] == 1
    e.close()
    # the steps follow the block tree: never more than the rows
    assert 1 <= st["fwd"]["steps"] <= c.n and 1 <= st["bwd"]["steps"] <= c.n


def test_level_schedule_steps_on_the_largest_cloth():
    """512 x 512 with shear and bending in the nested-dissection order: one step per
    level of the separator tree (18 per sweep), where chain pieces of 32 rows took
    368 forward and 140 backward steps."""
    c, _, _ = scenes.config3()
    e = Engine("mixed32", "fast")
    e.add(c, CFG)
    e.set_ordering("nd")
    e.set_launch("grid", 32)
    err = e.verify_schedule(0)
    st = e.schedule_stats(0)
    e.close()
    assert err <= 1e-12
    assert st["fwd"]["steps"] <= 20 and st["bwd"]["steps"] <= 20

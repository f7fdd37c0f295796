"""CPU-side pieces of bench.py: the instance-parallel CPU baseline runs oracle
instances concurrently on host threads, which is only a fair baseline if the
instances share no state.  These tests check that, on small cloths of the
512-instance mix."""
import numpy as np

import bench
from oracle_ffi import Oracle
from pd_ipc_amd import Engine, scenes


def _small_mix(count=4, max_n=1200):
    out = []
    k = 0
    while len(out) < count:
        c = scenes.config4_instance(4, k)
        if c.n <= max_n:
            out.append(c)
        k += 1
    cfg = scenes.SimConfig.fixed_iterations(40, scenes.H600)
    perms = []
    for c in out:
        e = Engine("fp64")
        e.add(c, cfg)
        perms.append(e.factor_arrays(0)["perm3"])
        e.close()
    return out, cfg, perms


def test_host_cores_positive():
    assert bench.host_cores() >= 1


def test_instance_parallel_counts():
    cloths, cfg, perms = _small_mix()
    inst, nsub, t = bench.cpu_baseline_instances(cloths, cfg, perms, 2, 3)
    assert inst == len(cloths) and nsub == 2 * len(cloths) and t > 0


def test_concurrent_oracles_match_serial_bits():
    from concurrent.futures import ThreadPoolExecutor
    cloths, cfg, perms = _small_mix()

    def run(k):
        o = Oracle(cloths[k], cfg, perms[k], threads=1)
        o.factor()
        o.step(2)
        return o.positions()

    serial = [run(k) for k in range(len(cloths))]
    with ThreadPoolExecutor(max_workers=len(cloths)) as pool:
        par = list(pool.map(run, range(len(cloths))))
    for a, b in zip(serial, par):
        assert np.array_equal(a, b)

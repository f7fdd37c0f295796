"""GPU adjoint (pdx_rollout / pdx_adjoint through the C ABI) against the FP64 CPU
adjoint of the oracle and against central finite differences of the GPU FP64
forward (step 1e-6), and the torch.autograd.Function wrapper over it."""
import numpy as np
import pytest

from pd_ipc_amd import Engine, SimConfig, scenes
from oracle_ffi import Oracle
from test_adjoint_cpu import Rollout, _scene, _traj

pytestmark = pytest.mark.gpu


def _cpu_reference(c, cfg, perm, nsub, traj, seed=9):
    r0 = Rollout(c, cfg, perm, nsub, np.zeros(3 * c.n))
    _, xT = r0.run(traj=traj)
    target = xT + 0.01 * np.random.default_rng(seed).standard_normal(xT.shape)
    r = Rollout(c, cfg, perm, nsub, target)
    return xT, target, r.adjoint(traj)


def _gpu(c, cfg, nsub, every, traj, target, precision, solve, ordering="amd"):
    e = Engine(precision, solve)
    e.add(c, cfg)
    if ordering != "amd":
        e.set_ordering(ordering)
    e.set_pin_trajectory(0, traj)
    e.rollout(nsub, every)
    xT, vT = e.state(0)
    e.set_adjoint_seed(0, 2.0 * (xT - target))
    e.adjoint()
    g = e.gradients(0)
    x2, v2 = e.state(0)
    assert np.array_equal(x2, xT) and np.array_equal(v2, vT)  # the adjoint leaves the state
    e.close()
    return xT, g


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


KEYS = ("x0", "v0", "kappa", "bend", "pin")


def test_adjoint_exact_fp64_matches_cpu_adjoint():
    """Exact FP64: the rollout's final state is the oracle's bit for bit and every
    gradient agrees with the FP64 CPU adjoint (segments of 2 substeps over 5)."""
    c, cfg, perm = _scene()
    nsub = 5
    traj = _traj(c, nsub, cfg.dt)
    xT_cpu, target, gc = _cpu_reference(c, cfg, perm, nsub, traj)
    xT, g = _gpu(c, cfg, nsub, 2, traj, target, "fp64", "exact")
    assert np.array_equal(xT, xT_cpu)
    for k in KEYS:
        err = _rel(g[k], gc[k])
        print(k, "max rel err", err, "bitwise", np.array_equal(g[k], gc[k]))
        assert err <= 1e-12, k


@pytest.mark.parametrize("precision,solve,tol,ordering", [
    ("fp64", "fast", 1e-9, "amd"), ("mixed32", "fast", 2e-4, "amd"), ("mixed32", "exact", 2e-4, "amd"),
    # the nested-dissection order against the CPU adjoint in the reference's AMD order
    ("fp64", "fast", 1e-9, "nd"), ("mixed32", "fast", 2e-4, "nd")])
def test_adjoint_forms_match_cpu_adjoint(precision, solve, tol, ordering):
    c, cfg, perm = _scene(nx=14, ny=11, iters=8)
    nsub = 7
    traj = _traj(c, nsub, cfg.dt)
    xT_cpu, target, gc = _cpu_reference(c, cfg, perm, nsub, traj)
    xT, g = _gpu(c, cfg, nsub, 3, traj, target, precision, solve, ordering)
    print(precision, solve, ordering, "max |xT - xT_fp64| =", float(np.max(np.abs(xT - xT_cpu))))
    for k in KEYS:
        err = _rel(g[k], gc[k])
        print(k, "max rel err vs FP64 CPU adjoint", err)
        assert err <= tol, k


def test_adjoint_batched_instances():
    """Uneven instances in one launch: each instance's gradients equal its own
    single-instance CPU adjoint."""
    cases = []
    for (nx, ny, it) in [(9, 7, 6), (13, 6, 6), (6, 12, 6)]:
        c, cfg, perm = _scene(nx=nx, ny=ny, iters=it)
        cases.append((c, cfg, perm))
    nsub = 4
    e = Engine("fp64", "fast")
    refs = []
    for c, cfg, perm in cases:
        traj = _traj(c, nsub, cfg.dt)
        xT_cpu, target, gc = _cpu_reference(c, cfg, perm, nsub, traj)
        refs.append((xT_cpu, target, gc))
        i = len(refs) - 1
        e.add(c, cfg)
        e.set_pin_trajectory(i, traj)
    e.rollout(nsub, 3)
    for i, (xT_cpu, target, gc) in enumerate(refs):
        xT, _ = e.state(i)
        assert np.max(np.abs(xT - xT_cpu)) <= 1e-12
        e.set_adjoint_seed(i, 2.0 * (xT - target))
    e.adjoint()
    for i, (_, _, gc) in enumerate(refs):
        g = e.gradients(i)
        for k in KEYS:
            assert _rel(g[k], gc[k]) <= 1e-9, (i, k)
    e.close()


def _gpu_loss(c, cfg, nsub, traj, target, kappa_scale=1.0, v0=None):
    e = Engine("fp64", "fast")
    cc = scenes.Cloth(**{**c.__dict__})
    cc.edge_kappa = c.edge_kappa * kappa_scale
    cc.bend_k = c.bend_k * kappa_scale
    if v0 is not None:
        cc.velocities = v0
    e.add(cc, cfg)
    e.set_pin_trajectory(0, traj)
    e.rollout(nsub, 5)
    x, _ = e.state(0)
    e.close()
    return float(np.sum((x - target) ** 2))


def test_adjoint_vs_central_differences_gpu():
    """48 x 40 cloth, 12 substeps, 10 iterations: directional derivatives of the
    GPU FP64 loss (central differences, step 1e-6) against the GPU adjoint."""
    top = 39 * 48
    c = scenes.grid_cloth(48, 40, 2300.0, shear=460.0, bend=46.0, pins=[top, top + 24, top + 47])
    c.velocities = 0.1 * np.random.default_rng(3).standard_normal(3 * c.n)
    cfg = SimConfig.fixed_iterations(10, 1.0 / 600.0)
    nsub = 12
    traj = _traj(c, nsub, cfg.dt)
    e = Engine("fp64", "fast")
    e.add(c, cfg)
    e.set_pin_trajectory(0, traj)
    e.rollout(nsub, 5)
    xT, _ = e.state(0)
    target = xT + 0.01 * np.random.default_rng(4).standard_normal(xT.shape)
    e.set_adjoint_seed(0, 2.0 * (xT - target))
    e.adjoint()
    g = e.gradients(0)
    e.close()
    eps = 1e-6
    # stiffness scale of the instance.  The FP64 forward carries ~1e-12 m of
    # rounding noise (second differences of x_T along the scale), which at a
    # 1e-6 step on this weakly sensitive parameter is ~1e-3 of the derivative;
    # the scale is therefore differenced at a relative step of 1e-3 (truncation
    # ~1e-6 relative, measured).
    ks = 1e-3
    fd = (_gpu_loss(c, cfg, nsub, traj, target, 1 + ks) - _gpu_loss(c, cfg, nsub, traj, target, 1 - ks)) / (2 * ks)
    ad = g["kappa"] @ c.edge_kappa + g["bend"] @ c.bend_k
    print("kappa scale: fd", fd, "adjoint", ad)
    assert abs(fd - ad) <= 1e-4 * abs(fd)
    # initial velocities along a seeded direction
    d = np.random.default_rng(5).standard_normal(3 * c.n)
    d[np.repeat(c.pinned.astype(bool), 3)] = 0.0
    fd = (_gpu_loss(c, cfg, nsub, traj, target, v0=c.velocities + eps * d) -
          _gpu_loss(c, cfg, nsub, traj, target, v0=c.velocities - eps * d)) / (2 * eps)
    print("v0: fd", fd, "adjoint", g["v0"] @ d)
    assert abs(fd - g["v0"] @ d) <= 1e-4 * abs(fd)
    # pin trajectory along a seeded direction
    dp = np.random.default_rng(6).standard_normal(traj.shape)
    fd = (_gpu_loss(c, cfg, nsub, traj + eps * dp, target) - _gpu_loss(c, cfg, nsub, traj - eps * dp, target)) / (2 * eps)
    ad = float(np.sum(g["pin"] * dp))
    print("pins: fd", fd, "adjoint", ad)
    assert abs(fd - ad) <= 1e-4 * abs(fd)


def test_torch_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from pd_ipc_amd.autograd import DifferentiableBatch, rollout
    c1, cfg, _ = _scene()
    c2, _, _ = _scene()
    c2.velocities = -c2.velocities
    nsub = 4
    batch = DifferentiableBatch([c1, c2], cfg, kappa_ref=[800.0, 800.0], nsub=nsub, checkpoint_every=2,
                                precision="fp64", solve="fast")
    traj = np.stack([_traj(c1, nsub, cfg.dt), _traj(c2, nsub, cfg.dt)])
    kappa = torch.tensor([800.0, 950.0], dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(np.stack([c1.velocities.reshape(-1, 3), c2.velocities.reshape(-1, 3)]), requires_grad=True)
    pins = torch.tensor(traj, requires_grad=True)
    target = torch.zeros(2, c1.n, 3, dtype=torch.float64)
    target[:, :, 1] = -0.01
    xT = rollout(kappa, v0, pins, batch)
    loss = ((xT - target) ** 2).sum()
    loss.backward()
    assert kappa.grad.shape == (2,) and v0.grad.shape == v0.shape and pins.grad.shape == pins.shape
    # central differences of the same wrapper's forward (stiffness at a relative
    # step of 1e-4, velocities at 1e-6)
    eps = 1e-6
    with torch.no_grad():
        for i in range(2):
            dk = torch.zeros(2, dtype=torch.float64)
            dk[i] = 800.0 * 1e-4
            lp = ((rollout(kappa + dk, v0, pins, batch) - target) ** 2).sum().item()
            lm = ((rollout(kappa - dk, v0, pins, batch) - target) ** 2).sum().item()
            fd = (lp - lm) / (2 * 800.0 * 1e-4)
            print("kappa", i, fd, kappa.grad[i].item())
            assert abs(fd - kappa.grad[i].item()) <= 1e-4 * abs(fd)
        d = torch.from_numpy(np.random.default_rng(7).standard_normal(tuple(v0.shape)))
        lp = ((rollout(kappa, v0 + eps * d, pins, batch) - target) ** 2).sum().item()
        lm = ((rollout(kappa, v0 - eps * d, pins, batch) - target) ** 2).sum().item()
        fd = (lp - lm) / (2 * eps)
        ad = float((v0.grad * d).sum())
        print("v0", fd, ad)
        assert abs(fd - ad) <= 1e-4 * abs(fd)

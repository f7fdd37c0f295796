"""torch.autograd.Function over the differentiable rollout of the C ABI.

A thin test wrapper: the forward runs `pdx_rollout` (fixed iterations,
checkpoints every k substeps) from fixed initial positions, the backward seeds
`pdx_set_adjoint_seed` with the incoming gradient of the final positions and
runs `pdx_adjoint`.  Tensors cross the boundary as host FP64 arrays; the state,
factor, tape and adjoint live on the GPU inside the engine.

Parameters per instance i:
  kappa[i]        stiffness scale: edge e gets kappa[i] * alpha_e, stencil s gets
                  kappa[i] * beta_s, with alpha / beta the ratios of the cloth's
                  own stiffnesses to its reference kappa (distance 1, shear and
                  bend their fractions);
  v0[i]           initial velocities (n, 3);
  pins[i]         pin targets per substep (nsub, pins, 3).
Output: final positions (ninst, n, 3).  All instances must have the same n and
pin count (the differentiable workload is a batch of equal grids).
"""
import numpy as np
import torch

from .pdx import Engine


class DifferentiableBatch:
    def __init__(self, cloths, cfg, kappa_ref, nsub, checkpoint_every=10, precision="mixed32",
                 solve="fast"):
        self.cloths = cloths
        self.nsub = nsub
        self.every = checkpoint_every
        self.engine = Engine(precision, solve)
        for c in cloths:
            self.engine.add(c, cfg)
        self.engine.factor()
        self.kappa_ref = np.asarray(kappa_ref, np.float64)
        self.alpha = [c.edge_kappa / k for c, k in zip(cloths, self.kappa_ref)]
        self.beta = [c.bend_k / k for c, k in zip(cloths, self.kappa_ref)]
        self.kappa = self.kappa_ref.copy()
        self.n = cloths[0].n
        self.npin = self.engine.pin_count(0)

    def _set_params(self, kappa, v0, pins):
        e = self.engine
        for i, c in enumerate(self.cloths):
            if kappa[i] != self.kappa[i]:
                e.set_edge_kappa(i, self.alpha[i] * kappa[i])
                if len(self.beta[i]):
                    e.set_bend_k(i, self.beta[i] * kappa[i])
            e.set_state(i, c.positions, v0[i].reshape(-1))
            e.set_pin_trajectory(i, pins[i])
        self.kappa = np.array(kappa, np.float64)

    def forward(self, kappa, v0, pins):
        self._set_params(kappa, v0, pins)
        self.engine.rollout(self.nsub, self.every)
        return np.stack([self.engine.state(i)[0].reshape(-1, 3) for i in range(len(self.cloths))])

    def backward(self, gx):
        e = self.engine
        for i in range(len(self.cloths)):
            e.set_adjoint_seed(i, gx[i].reshape(-1))
        e.adjoint()
        gk = np.empty(len(self.cloths))
        gv = np.empty((len(self.cloths), self.n, 3))
        gp = np.empty((len(self.cloths), self.nsub, self.npin, 3))
        for i in range(len(self.cloths)):
            g = e.gradients(i)
            gk[i] = g["kappa"] @ self.alpha[i] + (g["bend"] @ self.beta[i] if len(self.beta[i]) else 0.0)
            gv[i] = g["v0"].reshape(-1, 3)
            gp[i] = g["pin"]
        return gk, gv, gp


class RolloutFunction(torch.autograd.Function):
    """x_T = rollout(kappa, v0, pins); backward through the GPU adjoint.  One
    backward per forward: the adjoint differentiates the engine's last rollout."""

    @staticmethod
    def forward(ctx, kappa, v0, pins, batch):
        xT = batch.forward(kappa.detach().cpu().double().numpy(), v0.detach().cpu().double().numpy(),
                           pins.detach().cpu().double().numpy())
        ctx.batch = batch
        ctx.dev = kappa.device
        return torch.from_numpy(xT).to(kappa.device)

    @staticmethod
    def backward(ctx, gx):
        gk, gv, gp = ctx.batch.backward(gx.detach().cpu().double().numpy())
        dev = ctx.dev
        return torch.from_numpy(gk).to(dev), torch.from_numpy(gv).to(dev), torch.from_numpy(gp).to(dev), None


def rollout(kappa, v0, pins, batch):
    return RolloutFunction.apply(kappa, v0, pins, batch)

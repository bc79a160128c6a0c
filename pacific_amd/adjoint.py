"""Differentiable DEM rollouts: a ``torch.autograd.Function`` over the
discrete adjoint of ``libgrains_gpu.so`` (``grains_gpu_adjoint_forward`` /
``grains_gpu_adjoint_backward``).

    x_N, v_N, w_N = rollout(system, x0, v0, w0, theta, nsteps, checkpoint_every)

with ``theta = [kn, en, muc]`` of one material pair.  Gradients flow to x0,
v0, w0 ([n, 3], particle-id order) and theta.  The step is the FP64
SecondOrderLeapFrog step of spheres; torch only carries the arrays across
the C ABI (float64, any device; the library reads and writes host copies).
"""
import numpy as np
import torch

from .grains import GrainsGPU


def _np(t):
    return None if t is None else t.detach().to("cpu", torch.float64).numpy()


class DEMRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, w0, theta, system, nsteps, checkpoint_every, materials, etat):
        if not isinstance(system, GrainsGPU):
            raise TypeError("system must be a GrainsGPU")
        kn, en, muc = (float(a) for a in theta.detach().cpu().tolist())
        system.contact_model(materials[0], materials[1], kn, en, muc, kr=0.0, etat=etat)
        system.set_state(x=_np(x0), v=_np(v0), w=_np(w0), recompute=True)
        system.adjoint_forward(nsteps, checkpoint_every)
        st = system.state()
        ctx.system, ctx.materials, ctx.device = system, materials, x0.device
        return tuple(torch.from_numpy(st[k]).to(x0.device) for k in ("x", "v", "w"))

    @staticmethod
    def backward(ctx, gx, gv, gw):
        g = ctx.system.adjoint_backward(_np(gx), _np(gv), _np(gw), ctx.materials)
        dev = ctx.device
        out = [torch.from_numpy(g[k]).to(dev) for k in ("x0", "v0", "w0")]
        return out[0], out[1], out[2], torch.from_numpy(g["params"]).to(dev), None, None, None, None, None


def rollout(system, x0, v0, w0, theta, nsteps, checkpoint_every=50, materials=(0, 0), etat=-1.0):
    """nsteps differentiable steps of ``system`` from (x0, v0, w0) with the
    contact parameters theta = [kn, en, muc] of the material pair."""
    return DEMRollout.apply(x0, v0, w0, theta, system, int(nsteps), int(checkpoint_every), tuple(materials),
                            float(etat))

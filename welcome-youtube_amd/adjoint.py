"""torch.autograd wrapper of the window gradient.

    rdf = RDFWindow.apply(theta, engine, nsteps, spec, ckpt_every)
    loss = ((rdf - target) ** 2).sum()
    loss.backward()                       # theta.grad = dL/dtheta

theta is a tensor of shape (7,) or (nrep, 7) over (a_WW, a_WH, a_WT, a_HH,
a_HT, a_TT, gamma).  forward() sets those parameters on every replica and runs
the window on the GPU (oc_window_forward: the smooth tail-tail RDF averaged
over the window, checkpoints every ckpt_every steps); backward() receives
dL/drdf from torch and runs the hand-written adjoint sweep on the GPU
(oc_window_backward), with the noise stream fixed and chemistry frozen.
PyTorch only carries the 7 parameters and the RDF; the engine state stays in
the library's own device buffers.
"""
import numpy as np
import torch

from .abi import rdf_spec  # noqa: F401  (re-exported for callers)


class RDFWindow(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, engine, nsteps, spec, ckpt_every=20):
        th = theta.detach().to("cpu", torch.float64).numpy()
        if th.ndim == 1:
            th = np.broadcast_to(th, (engine.nrep, 7))
        for r in range(engine.nrep):
            engine.set_theta(th[r], r)
        g = engine.window_forward(nsteps, spec, ckpt_every)
        ctx.engine = engine
        ctx.shape = tuple(theta.shape)
        ctx.device, ctx.dtype = theta.device, theta.dtype
        out = torch.from_numpy(g)
        return out[0] if len(ctx.shape) == 1 and engine.nrep == 1 else out

    @staticmethod
    def backward(ctx, grad_out):
        go = grad_out.detach().to("cpu", torch.float64).numpy()
        if go.ndim == 1:
            go = go[None, :]
        g = ctx.engine.window_backward(go)
        gt = torch.from_numpy(g)
        if len(ctx.shape) == 1:
            gt = gt.sum(0)
        return gt.to(ctx.device, ctx.dtype), None, None, None, None


def window_loss(theta, engine, nsteps, spec, target, ckpt_every=20):
    """L2 distance of the window-mean RDF to target (summed over replicas)."""
    rdf = RDFWindow.apply(theta, engine, nsteps, spec, ckpt_every)
    t = torch.as_tensor(np.asarray(target), dtype=rdf.dtype)
    return ((rdf - t) ** 2).sum()

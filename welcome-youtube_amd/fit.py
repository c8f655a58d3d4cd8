"""Batched parameter fit: Adam on the repulsion matrix toward a target RDF.

    hist = fit_rdf(engine, spec, target, steps=50, lr=0.5)

Every optimiser step runs one adjoint window on all replicas of one engine
(oc_window_forward, then oc_window_backward through RDFWindow): the smooth
tail-tail RDF of each replica averaged over the window, the loss
L = (1/R) sum_r sum_b (gbar_rb - target_b)^2 and its gradient with respect to
one theta shared by the R replicas, which is the mean of the per-replica
adjoint gradients.  torch.optim.Adam then updates the free entries of theta.
With rewind=True every window starts from the replicas' state at the first
step (oc_window_rewind): the loss is then one deterministic function of theta
(fixed start, fixed noise), at the same age of the trajectories as a target
taken at that age.  With rewind=False the replicas continue from each window's
last state, so successive windows sample successive stretches of each
trajectory.

theta = (a_WW, a_WH, a_WT, a_HH, a_HT, a_TT, gamma); `free` selects the
entries Adam moves (default: the six a_ij, gamma fixed).  The noise stream is
fixed and chemistry is off inside each window, as for a single window
gradient.  The reference has no counterpart (its parameters are fixed for a
run).
"""
import time

import numpy as np
import torch

from .adjoint import RDFWindow

A_IJ = (0, 1, 2, 3, 4, 5)


def fit_rdf(engine, spec, target, steps=50, lr=0.5, window=200, ckpt_every=20, free=A_IJ, theta0=None,
            betas=(0.9, 0.999), rewind=True, callback=None):
    """Adam steps on the shared theta of every replica of `engine`.

    Returns a list with one dict per step: the loss and gradient of the window
    run at that step's theta (before the update), the theta it ran at, and the
    wall time of the window (forward + adjoint sweep) and of the update."""
    free = tuple(int(q) for q in free)
    if not free or min(free) < 0 or max(free) > 6 or len(set(free)) != len(free):
        raise ValueError(f"free must list distinct entries of 0..6, got {free}")
    t = np.asarray(target, dtype=np.float64)
    if t.shape != (spec.nbins,):
        raise ValueError(f"target must have {spec.nbins} bins, got shape {t.shape}")
    th0 = np.asarray(engine.theta(0) if theta0 is None else theta0, dtype=np.float64)
    if th0.shape != (7,):
        raise ValueError("theta0 must have 7 entries")
    base = torch.tensor(th0, dtype=torch.float64)
    idx = torch.tensor(free, dtype=torch.long)
    x = base[idx].clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=lr, betas=betas)
    tgt = torch.as_tensor(t)
    R = engine.nrep
    hist = []
    for k in range(steps):
        opt.zero_grad()
        theta = base.index_put((idx,), x)
        t0 = time.perf_counter()
        rdf = RDFWindow.apply(theta.expand(R, 7), engine, window, spec, ckpt_every)
        loss = ((rdf.reshape(R, -1) - tgt) ** 2).sum() / R
        loss.backward()
        t1 = time.perf_counter()
        g = np.zeros(7)
        g[list(free)] = x.grad.numpy()
        rec = {"step": k, "loss": float(loss.item()), "theta": theta.detach().numpy().copy(), "grad": g,
               "window_s": t1 - t0}
        opt.step()
        if rewind:
            engine.window_rewind()
        rec["update_s"] = time.perf_counter() - t1
        hist.append(rec)
        if callback is not None:
            callback(rec)
    final = base.index_put((idx,), x.detach()).numpy()
    for r in range(R):
        engine.set_theta(final, r)
    return hist

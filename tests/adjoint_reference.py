"""Dense FP64 CPU autograd reference for the differentiable LJ step.

Test infrastructure: the checker for the hand-written adjoint kernels.

Potential: V(r; eps, sigma) = eps * v(r / sigma) with the reference's
quadratically shifted form v(x) = 4 (x^-12 - x^-6 + C2 x^2 + C0)
(variables.cc:21-27) cut at x = 3, so V(3 sigma) = V'(3 sigma) = 0; at
eps = sigma = 1 this is exactly the reference's potential.  Dynamics:
drift-kick-drift (mdmanager.cc:154-186).  Loss over the last `nobs` of
`nsteps` steps, observed after each full step:

    L = (mean_k P_k - p_target)^2 + (mean_k PE_k / N - pe_target)^2,
    P_k = (2 KE_k / 3 - W_k) / V,  W = sum_pairs r V'(r) / 3

(pressure as MDManager::Pressure, mdmanager.cc:563-570).  Pairs are every
minimum-image pair of the box (dense N^2), no neighbour list.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def shift_constants(rc=3.0):
    c2 = 6.0 * rc ** -14 - 3.0 * rc ** -8
    c0 = -rc ** -12 + rc ** -6 - c2 * rc * rc
    return c0, c2


C0, C2 = shift_constants()


def pair_terms(q, L, eps, sigma):
    """(pair energy sum, pair virial sum W, forces) with dense min image."""
    d = q[None, :, :] - q[:, None, :]                       # d_ij = q_j - q_i
    d = d - L * torch.round(d / L)
    s = (d * d).sum(-1)
    n = q.shape[0]
    iu = torch.triu_indices(n, n, 1)
    s_p = s[iu[0], iu[1]]
    x2 = s_p / (sigma * sigma)
    mask = x2 <= 9.0
    x2m = torch.where(mask, x2, torch.ones_like(x2))
    inv2 = 1.0 / x2m
    inv6 = inv2 ** 3
    v = 4.0 * (inv6 * inv6 - inv6 + C2 * x2m + C0)
    # r V'(r) = eps * x v'(x) with x v'(x) = 4(-12 x^-12 + 6 x^-6 + 2 C2 x^2)
    xdv = 4.0 * (-12.0 * inv6 * inv6 + 6.0 * inv6 + 2.0 * C2 * x2m)
    pe = eps * torch.where(mask, v, torch.zeros_like(v)).sum()
    W = eps * torch.where(mask, xdv, torch.zeros_like(xdv)).sum() / 3.0
    return pe, W


def forces(q, L, eps, sigma, create_graph):
    pe, _ = pair_terms(q, L, eps, sigma)
    (g,) = torch.autograd.grad(pe, q, create_graph=create_graph)
    return -g


def step(q, p, L, eps, sigma, dt, create_graph):
    qh = q + p * (dt * 0.5)
    p1 = p + forces(qh, L, eps, sigma, create_graph) * dt
    q1 = qh + p1 * (dt * 0.5)
    return q1, p1


def loss_and_grads(q0, p0, L, eps=1.0, sigma=1.0, dt=0.001, nsteps=200, nobs=100,
                   p_target=1.5, pe_target=-4.0, segment=20):
    """Loss and dL/d(eps, sigma, q0) by autograd, recomputing per segment."""
    dtype = torch.float64
    Lt = torch.tensor(np.asarray(L), dtype=dtype)
    V = float(np.prod(L))
    n = q0.shape[0]
    e = torch.tensor(eps, dtype=dtype, requires_grad=True)
    sg = torch.tensor(sigma, dtype=dtype, requires_grad=True)
    q = torch.tensor(q0, dtype=dtype, requires_grad=True)
    p = torch.tensor(p0, dtype=dtype)
    qstart = q

    # forward, keeping only segment starts (values) for the observables
    obs_P, obs_E = [], []

    def run_segment(q, p, k0, k1, keep_graph):
        Ps, Es = [], []
        for k in range(k0, k1):
            q, p = step(q, p, Lt, e, sg, dt, create_graph=keep_graph)
            if k >= nsteps - nobs:
                pe, W = pair_terms(q, Lt, e, sg)
                ke = 0.5 * (p * p).sum()
                Ps.append((2.0 * ke / 3.0 - W) / V)
                Es.append(pe / n)
        return q, p, Ps, Es

    # full autograd graph, checkpointed by torch.utils.checkpoint per segment
    from torch.utils.checkpoint import checkpoint
    Psum = torch.zeros((), dtype=dtype)
    Esum = torch.zeros((), dtype=dtype)
    for k0 in range(0, nsteps, segment):
        k1 = min(nsteps, k0 + segment)

        def seg(qq, pp, ee, ss, k0=k0, k1=k1):
            Ps, Es = [], []
            for k in range(k0, k1):
                qh = qq + pp * (dt * 0.5)
                pe_h, _ = pair_terms(qh, Lt, ee, ss)
                (gq,) = torch.autograd.grad(pe_h, qh, create_graph=True)
                pp = pp - gq * dt
                qq = qh + pp * (dt * 0.5)
                if k >= nsteps - nobs:
                    pe, W = pair_terms(qq, Lt, ee, ss)
                    ke = 0.5 * (pp * pp).sum()
                    Ps.append((2.0 * ke / 3.0 - W) / V)
                    Es.append(pe / n)
            zero = torch.zeros((), dtype=dtype)
            return qq, pp, (torch.stack(Ps).sum() if Ps else zero), (torch.stack(Es).sum() if Es else zero)

        if not p.requires_grad:
            p = p.detach().requires_grad_(False)
        q, p, ps, es = checkpoint(seg, q, p, e, sg, use_reentrant=False)
        Psum = Psum + ps
        Esum = Esum + es
    Pm = Psum / nobs
    Em = Esum / nobs
    loss = (Pm - p_target) ** 2 + (Em - pe_target) ** 2
    gq, ge, gs = torch.autograd.grad(loss, (qstart, e, sg))
    return (float(loss.detach()), float(ge), float(gs), gq.numpy(), float(Pm.detach()),
            float(Em.detach()))


def loss_only(q0, p0, L, eps=1.0, sigma=1.0, dt=0.001, nsteps=200, nobs=100,
              p_target=1.5, pe_target=-4.0):
    dtype = torch.float64
    Lt = torch.tensor(np.asarray(L), dtype=dtype)
    V = float(np.prod(L))
    n = q0.shape[0]
    with torch.no_grad():
        q = torch.tensor(q0, dtype=dtype)
        p = torch.tensor(p0, dtype=dtype)
    Ps, Es = [], []
    for k in range(nsteps):
        qh = (q + p * (dt * 0.5)).detach().requires_grad_(True)
        pe_h, _ = pair_terms(qh, Lt, eps, sigma)
        (gq,) = torch.autograd.grad(pe_h, qh)
        with torch.no_grad():
            p = p - gq * dt
            q = qh + p * (dt * 0.5)
            if k >= nsteps - nobs:
                pe, W = pair_terms(q, Lt, eps, sigma)
                ke = 0.5 * (p * p).sum()
                Ps.append(float((2.0 * ke / 3.0 - W) / V))
                Es.append(float(pe / n))
    Pm, Em = float(np.mean(Ps)), float(np.mean(Es))
    return (Pm - p_target) ** 2 + (Em - pe_target) ** 2, Pm, Em

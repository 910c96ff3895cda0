"""torch.autograd front end of the batched contact path.

contact_wrench(batch, E1, E2, phi1, phi2, kn, kt) returns the per-instance wrench
fm (n x 12, [tau1 f1 tau2 f2], each body's frame) of ForceGeneralSDFContact for the
given poses, twists and stiffnesses, and differentiates it with the library's
hand-written adjoint kernel (atac_batch_vjp):

* phi1, phi2 (n x 6), kn, kt (n): plain gradients.
* E1, E2 (n x 4 x 4 rigid transforms): the pose enters through the reference's
  perturbation E <- E exp([w; v]) (ForceGeneralSDFContact.cpp:1066), whose adjoint
  is the body-frame twist gradient g = (g_w, g_v) (atac_batch_vjp's g_q1 / g_q2).
  The gradient returned for E is the 4 x 4 matrix G = E^-T M(g) with
  M(g) = [[skew(g_w)/2, g_v], [0, 0]]: for every tangent direction dE = E xi^ it
  satisfies <G, dE> = g . xi, which is what any loss of the form L(E exp(xi))
  sees. Components of G orthogonal to the tangent space are zero in that pairing.

The forward pass uploads the poses (set_instances, a host copy) and evaluates; the
backward pass needs the records of that evaluation, so evaluating the batch again
between the two raises.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L

__all__ = ["contact_wrench", "ContactWrench", "se3_exp", "tangent_pairing_matrix"]


def se3_exp(xi: torch.Tensor) -> torch.Tensor:
    """exp of body twists xi = [w; v] (n x 6, FP64) as n x 4 x 4 transforms
    (the closed form of Utils.h:126-159)."""
    xi = xi.to(torch.float64)
    w, v = xi[:, :3], xi[:, 3:]
    th = torch.linalg.norm(w, dim=1)
    W = torch.zeros((xi.shape[0], 3, 3), dtype=torch.float64, device=xi.device)
    W[:, 0, 1], W[:, 0, 2], W[:, 1, 0] = -w[:, 2], w[:, 1], w[:, 2]
    W[:, 1, 2], W[:, 2, 0], W[:, 2, 1] = -w[:, 0], -w[:, 1], w[:, 0]
    W2 = W @ W
    small = th < 1e-8
    ths = torch.where(small, torch.ones_like(th), th)
    a = torch.where(small, 1.0 - th ** 2 / 6.0, torch.sin(ths) / ths)
    b = torch.where(small, 0.5 - th ** 2 / 24.0, (1.0 - torch.cos(ths)) / ths ** 2)
    c = torch.where(small, 1.0 / 6.0 - th ** 2 / 120.0, (ths - torch.sin(ths)) / ths ** 3)
    eye = torch.eye(3, dtype=torch.float64, device=xi.device).expand_as(W)
    R = eye + a[:, None, None] * W + b[:, None, None] * W2
    V = eye + b[:, None, None] * W + c[:, None, None] * W2
    out = torch.zeros((xi.shape[0], 4, 4), dtype=torch.float64, device=xi.device)
    out[:, :3, :3] = R
    out[:, :3, 3] = (V @ v[:, :, None])[:, :, 0]
    out[:, 3, 3] = 1.0
    return out


def tangent_pairing_matrix(E: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
    """G = E^-T M(g), M(g) = [[skew(g_w)/2, g_v], [0, 0]] (see the module docstring)."""
    gw, gv = g[:, :3], g[:, 3:]
    M = torch.zeros((g.shape[0], 4, 4), dtype=torch.float64, device=g.device)
    M[:, 0, 1], M[:, 0, 2], M[:, 1, 2] = -gw[:, 2] / 2, gw[:, 1] / 2, -gw[:, 0] / 2
    M[:, 1, 0], M[:, 2, 0], M[:, 2, 1] = gw[:, 2] / 2, -gw[:, 1] / 2, gw[:, 0] / 2
    M[:, :3, 3] = gv
    return torch.linalg.inv(E.to(torch.float64)).transpose(1, 2) @ M


class ContactWrench(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, E1, E2, phi1, phi2, kn, kt, precise):
        n = batch.n
        host = [t.detach().to("cpu", torch.float64).contiguous().numpy() for t in (E1, E2, phi1, phi2, kn, kt)]
        insts = []
        for i in range(n):
            s = L.Instance()
            C_copy = batch._inst[i]
            for name, _ in L.Instance._fields_:
                setattr(s, name, getattr(C_copy, name))
            s.E1[:] = host[0][i].ravel().tolist()
            s.E2[:] = host[1][i].ravel().tolist()
            s.phi1[:] = host[2][i].tolist()
            s.phi2[:] = host[3][i].tolist()
            s.kn = float(host[4][i])
            s.kt = float(host[5][i])
            insts.append(s)
        batch.set_instances(insts)
        out = batch.eval(precise=bool(precise))
        ctx.batch = batch
        ctx.generation = batch.generation
        ctx.save_for_backward(E1, E2)
        return out["fm"]

    @staticmethod
    def backward(ctx, g_fm):
        batch = ctx.batch
        if batch.generation != ctx.generation:
            raise RuntimeError("contact_wrench: the batch was evaluated again or given new instances after this forward pass; "
                               "its adjoint records are gone")
        E1, E2 = ctx.saved_tensors
        g = batch.vjp(g_fm.contiguous())
        gE1 = tangent_pairing_matrix(E1, g["g_q1"]) if ctx.needs_input_grad[1] else None
        gE2 = tangent_pairing_matrix(E2, g["g_q2"]) if ctx.needs_input_grad[2] else None
        return (None, gE1, gE2, g["g_phi1"], g["g_phi2"], g["g_kn"], g["g_kt"], None)


def contact_wrench(batch, E1, E2, phi1, phi2, kn, kt, precise=False):
    """Differentiable per-instance wrench of a ContactBatch (see the module docstring).
    All tensors on the batch's device, FP64: E1, E2 (n, 4, 4), phi1, phi2 (n, 6), kn, kt (n,)."""
    return ContactWrench.apply(batch, E1, E2, phi1, phi2, kn, kt, precise)

"""Eigenvalue gradients with respect to per-vertex permittivity and chirality.

A forward/backward pair over the C-ABI (no reference counterpart: the reference does
not differentiate its eigenvalues).  The pencil is the JD-SIRA one, fncB x = w fncA x
(Lanczos' ew = 1/w, fun_Lanczos/Lanczos.m), with fncB built from a per-vertex
bi-isotropic material (``Plan.set_vertex_material``).  For an eigenpair,

    dw/dp = x^H (d fncB / dp) x / x^H fncA x,

and x^H (d fncB/dp) x per vertex comes from ``fame_material_gradient`` (the adjoint of
the material stencil).  For a cluster of equal eigenvalues the individual derivatives
are not defined; the derivative of the cluster's mean is returned for each member
(Lanczos' eigenvectors are fncB-orthonormal, hence fncA-orthogonal, so the trace
formula applies to the cluster).
"""
from __future__ import annotations

import numpy as np


def clusters(omega, rtol=1e-8):
    """Index groups of (nearly) equal eigenvalues in one k-point's sorted omega."""
    groups, cur = [], [0]
    for j in range(1, len(omega)):
        if abs(omega[j] - omega[cur[0]]) <= rtol * max(abs(omega[cur[0]]), 1e-300):
            cur.append(j)
        else:
            groups.append(cur)
            cur = [j]
    groups.append(cur)
    return groups


def eigen_gradients(plan, ev, omega=None, cluster_rtol=1e-8):
    """dw_j/d eps_v and dw_j/d gamma_v for eigenvectors ev ((nb, k, 4N) device tensor).

    Returns (d_eps, d_gamma), device tensors (nb, k, N), with cluster means applied when
    ``omega`` ((nb, k), ascending per row) is given."""
    t = plan.torch
    x = ev.to(device=plan.device, dtype=t.complex128).contiguous()
    if x.dim() == 2:
        x = x.unsqueeze(0)
    nb, k, L = x.shape
    g = plan.material_gradient(x)            # (nb, k, 3, N)
    den = t.empty((nb, k), dtype=t.float64, device=plan.device)
    for j in range(k):
        a = plan.fncA(x[:, j].contiguous()).reshape(nb, L)
        den[:, j] = t.sum(t.conj(x[:, j]) * a, dim=1).real
    d_eps = g[:, :, 0] / den[:, :, None]
    d_gam = g[:, :, 1] / den[:, :, None]
    if omega is not None:
        om = np.asarray(omega, float).reshape(nb, k)
        for b in range(nb):
            for grp in clusters(om[b], cluster_rtol):
                if len(grp) > 1:
                    d_eps[b, grp] = d_eps[b, grp].mean(dim=0, keepdim=True)
                    d_gam[b, grp] = d_gam[b, grp].mean(dim=0, keepdim=True)
    return d_eps, d_gam


class BandGradient:
    """forward: the k smallest positive omega of each k-point (Lanczos on the device);
    backward: dL/d omega -> dL/d eps_v (per vertex) and dL/d gamma (per vertex)."""

    def __init__(self, plan):
        self.plan = plan
        self.ev = None
        self.omega = None

    def forward(self, k, **lanczos_kw):
        ev, ew, _ = self.plan.Lanczos(k, **lanczos_kw)
        omega = 1.0 / ew
        order = np.argsort(omega, axis=1)
        t = self.plan.torch
        idx = t.as_tensor(order, device=self.plan.device)
        self.ev = t.gather(ev, 1, idx[:, :, None].expand(-1, -1, ev.shape[2]))
        self.omega = np.take_along_axis(omega, order, axis=1)
        return self.omega

    def backward(self, dL_domega):
        """dL/d eps (nb, N) and dL/d gamma_v (nb, N) for dL/d omega of shape (nb, k)."""
        t = self.plan.torch
        d_eps, d_gam = eigen_gradients(self.plan, self.ev, self.omega)
        w = t.as_tensor(np.asarray(dL_domega, float), device=self.plan.device)
        return t.einsum("bk,bkn->bn", w, d_eps), t.einsum("bk,bkn->bn", w, d_gam)

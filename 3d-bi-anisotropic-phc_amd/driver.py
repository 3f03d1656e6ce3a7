"""Band-structure driver on the device: the per-k loop of
band_structure_woodpile_Lanczos_n60_gamma0p2.m:40-123 (SVD tables + material, Lanczos,
omega = 1/ew, eigenvector restoration, residuals, the .mat file), either one k-point at a
time (the reference's loop) or with all k-points of the path batched in one plan.
"""
from __future__ import annotations

import time

import numpy as np


def band_structure_lanczos(pkg, dims, thetas, table, ids, nev=10, tol=1e-10, itmax=10000, p=None, batched=True,
                           restore=True, mat_path=None, delta=None, cg_tol=1e-12, cg_maxit=10000):
    """Returns a dict with Freq_mtx (nev, nk), iter_records (nk,), compute_eigs (nk,) seconds per
    k-point (table setup through residuals, as the reference's time_eig window), not_converged
    (the k indices whose Lanczos stopped at itmax or whose inner CG reached cg_maxit; their
    columns hold the last Ritz values, as Lanczos.m returns at mm == maxiter, and their
    iter_records are negated), and, with ``restore``, the restored fields and the NFSEP /
    original-problem residuals."""
    import torch

    thetas = np.atleast_2d(np.asarray(thetas, float))
    nk = thetas.shape[0]
    p = 2 * nev if p is None else p
    lz = dict(tol=tol, maxiter=itmax, p=p, cg_tol=cg_tol, cg_maxit=cg_maxit, raise_on_maxit=False)
    F = np.zeros((nev, nk))
    iters = np.zeros(nk, int)
    secs = np.zeros(nk)
    e_h_all, rn_all, ro_all = [], [], []

    def finish(plan, ev, ew):
        omega = 1.0 / ew
        order = np.argsort(omega, axis=1)
        omega = np.take_along_axis(omega, order, axis=1)
        idx = torch.as_tensor(order, device=plan.device)
        ev = torch.gather(ev, 1, idx[:, :, None].expand(-1, -1, ev.shape[2]))
        if not restore:
            return omega, None, None, None
        e_h = plan.Eigen_Restoration_Chiral_no_inv(ev)
        rn, ro = plan.EvaluteError_Chiral_no_inv(ev, omega, e_h)
        return omega, e_h, rn, ro

    if batched:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan = pkg.Plan(dims, thetas, table, ids, delta=delta)
        plan.autotune()
        ev, ew, it = plan.Lanczos(nev, **lz)
        omega, e_h, rn, ro = finish(plan, ev, ew)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        F[:, :] = omega.T
        iters[:] = it
        secs[:] = dt / nk
        if restore:
            e_h_all = list(e_h.cpu().numpy())
            rn_all, ro_all = list(rn), list(ro)
        plan.close()
    else:
        for i in range(nk):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan = pkg.Plan(dims, thetas[i], table, ids, delta=delta)
            ev, ew, it = plan.Lanczos(nev, **lz)
            omega, e_h, rn, ro = finish(plan, ev, ew)
            torch.cuda.synchronize()
            secs[i] = time.perf_counter() - t0
            F[:, i] = omega[0]
            iters[i] = it[0]
            if restore:
                e_h_all.append(e_h.cpu().numpy()[0])
                rn_all.append(rn[0])
                ro_all.append(ro[0])
            plan.close()
    out = {"Freq_mtx": F, "iter_records": iters, "compute_eigs": secs,
           "not_converged": np.nonzero(iters < 0)[0].tolist()}
    if restore:
        out.update(e_h=np.stack(e_h_all), res_NFSEP=np.stack(rn_all), res_ORGSEP=np.stack(ro_all))
        if mat_path:
            pkg.outputs.save_band_structure(mat_path, F, out["e_h"], out["res_NFSEP"], out["res_ORGSEP"], secs, iters)
    return out

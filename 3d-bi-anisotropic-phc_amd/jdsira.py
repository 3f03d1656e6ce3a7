"""JD-SIRA with locking for  A x = lambda B x  (A, B Hermitian) over the device operators.

Host-side restatement of the reference's Jacobi-Davidson / shift-invert residual
Arnoldi driver:

* ``GEP_AB_Herm_JDSIRA_Driver4``   light_locked_Huang/GEP_AB_Herm_JDSIRA_Driver4.m:1-81
* ``_locking``                     light_locked_Huang/GEP_AB_Herm_JDSIRA_Locking4.m:1-507
* ``choose_ritz_pair``             light_locked_Huang/Choose_Ritz_Pair_Unsymm4.m:1-412
* ``gram_schmidt``                 light_locked_Huang/Gram_Schmidt.m:1-25

As in the reference drivers (main_EIG_np_sphere_n60_gamma3p46427.m:124), A is
``fncB`` and B is ``fncA``, so the eigenvalues are the frequencies omega.  Every
vector operation (operator applications, projections V^H w, V y, the MINRES inner
solve) runs on the GPU through the C-ABI; only the small projected matrices
(size <= no_restart + 2) live on the host, as the reference gathers them for
``eig`` (Locking4.m:115).  The projected pencil has M = V^H A V Hermitian positive
definite (A = fncB), so ``eig(M, L)`` is solved as the Hermitian-definite problem
L y = mu M y with lambda = 1/mu.

Deviation: when the second or a later Ritz pair has converged as well, the reference
locks ``init_Vec(:, j1+1)`` with a counter that only advances on locking
(Locking4.m:459-476), which pairs the wrong vector with the value whenever an
earlier candidate did not converge; here the i-th Ritz vector is locked with the
i-th Ritz value.

``sigma_inner``: the reference's inner operator ignores the shift
(Jacobi_Davidson/GEP_shift_mtx_prod_vec.m:28 applies A only); ``sigma_inner=None``
reproduces that, a number solves with A - sigma_inner*B (shift-invert variant).
"""
from __future__ import annotations

import contextlib

import numpy as np
import scipy.linalg

from .fame import FAME_ERR_NOT_CONVERGED, NotConverged


def _one_blas_thread():
    """The host work of an outer step is on matrices of at most no_restart + 2 rows, where a
    multithreaded BLAS spends more time starting threads than computing (a 52 x 52 Hermitian
    pencil: 0.6 ms on one thread, 1.2 ms on eight in one measurement)."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # the limit is a speed measure only
        return contextlib.nullcontext()
    return threadpool_limits(limits=1, user_api="blas")


class JDResult:
    def __init__(self, ew, ev, log, total_inner, epiter):
        self.ew = ew            # converged eigenvalues (host, complex)
        self.ev = ev            # device tensor (ew_number, len)
        self.log = log          # list of (ew_order, EPIteNo, ritz value, residual) lines
        self.total_inner = total_inner
        self.outer_iterations = epiter


def gram_schmidt(plan, Y):
    """Modified Gram-Schmidt, dropping vectors of norm <= 1e-6 (Gram_Schmidt.m:1-25).

    Y is a device tensor (n_Y, len) modified in place; returns the number kept.  In the
    inner loop (Gram_Schmidt.m:13-16) each Y_j, j > k, meets only the new Y_nQ, so the
    coefficients Y_nQ^H Y_j come from one block product and stay on the device."""
    n_Y = Y.shape[0]
    n_Q = 0
    for k in range(n_Y):
        temp = _norm(plan, Y[k])
        if abs(temp) > 1.0e-6:
            Y[n_Q] = Y[k] / temp
            if k + 1 < n_Y:
                c = -plan.block_dot(Y[k + 1:], Y[n_Q]).conj()     # -(Y_nQ^H Y_j)
                q = Y[n_Q].reshape(1, -1)
                for j in range(k + 1, n_Y):
                    plan.block_gemv(q, c[j - k - 1:j - k], out=Y[j], beta=1.0)
            n_Q += 1
    return n_Q


def _dot(plan, x, y):
    return complex(plan.block_dot(x.reshape(1, -1), y).cpu().numpy()[0])


def _norm(plan, x):
    return float(np.sqrt(abs(_dot(plan, x, x))))


def _norm_dev(plan, x):
    """||x|| as a 0-dim device tensor: for normalisations that need no host decision, so the
    stream is not drained for a round trip."""
    return plan.block_dot(x.reshape(1, -1), x)[0].abs().sqrt()


def choose_ritz_pair(iteno, ew, ew_order, target, VRR, vec_r_tol, choose_type, rppd, target_org, choose_type_org,
                     zflag_ew, flag_no_ew):
    """Select Ritz pairs (Choose_Ritz_Pair_Unsymm4.m:1-412); returns (ewno, ritz_ew, ritz_ev).

    The matrices are at most a few tens wide, so the scans run over Python lists (numpy's
    per-call overhead exceeds the work); the close-value tests come from one distance matrix."""
    ew = np.asarray(ew, complex)
    ewl = ew[:iteno].tolist()
    sort_ritz = list(ewl)
    sort_No = 5
    # The reference tests closeness with a loop k = -1, 0, ..., K that stops at the first
    # |a - b| <= 10^k * vec_r_tol; the thresholds grow with k, so the loop succeeds exactly
    # when |a - b| <= 10^K * vec_r_tol (K = 1 for multiplicities, sort_No for the removal).
    tol_mult = 10.0 * vec_r_tol
    tol_remove = 10.0 ** sort_No * vec_r_tol
    if ew_order > 1:
        conv = np.asarray(zflag_ew[:flag_no_ew], complex)
        convl = conv.tolist()
        live = [True] * flag_no_ew            # sort_ew(i) > -1e13 in the reference
        mult_cn_ew = [0] * flag_no_ew
        cc = np.abs(conv[:, None] - conv[None, :]) <= tol_mult
        for i in range(flag_no_ew):  # multiplicity of converged eigenvalues (:52-78)
            if live[i]:
                no_cn_ew = 1
                for i1 in np.nonzero(cc[i, i + 1:])[0].tolist():
                    i1 += i + 1
                    if live[i1]:
                        no_cn_ew += 1
                        live[i1] = False
            else:
                no_cn_ew = 0
            mult_cn_ew[i] = no_cn_ew
        idx_ritz_w = [0] * iteno  # 1-based links as in the reference (0 = first occurrence)
        mult_ritz_w = [0] * iteno
        ea = ew[:iteno]
        close = np.abs(ea[:, None] - ea[None, :]) <= tol_mult
        partners = {}
        for i, j in np.argwhere(np.triu(close, 1)).tolist():
            partners.setdefault(i, []).append(j)
        for i in range(iteno):  # multiplicity of Ritz values (:97-121)
            if idx_ritz_w[i] == 0:
                # the inner loop only links indices it has not passed yet, so its set of
                # links is fixed by the state before it
                js = [j for j in partners.get(i, ()) if idx_ritz_w[j] == 0]
                for j in js:
                    idx_ritz_w[j] = i + 1
                mult_ritz_w[i] = 1 + len(js)
            else:
                mult_ritz_w[i] = 0
        near = np.abs(conv[:, None] - ea[None, :]) <= tol_remove
        removed = [False] * iteno  # sort_ritz set to 1e14 (never within tol_remove again)
        for i in range(flag_no_ew):  # remove Ritz values equal to converged eigenvalues (:126-228)
            if not live[i]:
                continue
            ind_sort = -1
            for r in np.nonzero(near[i])[0].tolist():
                if idx_ritz_w[r] == 0 and not removed[r]:
                    ind_sort = r
                    break
            if ind_sort < 0:
                continue  # 'Error in Choose_Ritz_Pair' (printed by the reference, no action)
            linked = [jj for jj in range(iteno) if idx_ritz_w[jj] == ind_sort + 1]
            if mult_cn_ew[i] == mult_ritz_w[ind_sort] or mult_cn_ew[i] > mult_ritz_w[ind_sort]:
                for r in [ind_sort] + linked:
                    sort_ritz[r] = 1.0e14
                    removed[r] = True
            else:
                idx_mult = [ind_sort] + linked
                ld = [float(np.linalg.norm(VRR[ew_order - 1:iteno, r])) for r in idx_mult]
                for _ in range(mult_cn_ew[i]):
                    kmin = 0
                    tmp = ld[0]
                    for i2 in range(1, mult_ritz_w[ind_sort]):
                        if tmp > ld[i2]:
                            tmp, kmin = ld[i2], i2
                    sort_ritz[idx_mult[kmin]] = 1.0e14
                    removed[idx_mult[kmin]] = True
                    ld[kmin] = 1.0e14
            idx_ritz_w[ind_sort] = -1
    # candidates (:235-306)
    tr = float(np.real(target))
    ewp_idx, dist = [], []
    for i in range(iteno):
        if abs(sort_ritz[i]) < 1.0e10:
            e = ewl[i]
            if choose_type == "CL":
                ok = (abs(e.real) < 1e-4 and e.imag > 1e-3) or (e.real > tr and abs(e.imag) < 1e-8)
                d = abs(e.real - tr)
            elif choose_type == "RGTR":
                ok = e.real > tr and abs(e.imag) < 1.0e-12
                d = abs(e.real - tr)
            elif choose_type == "RGTC":
                ok = e.real > tr
                d = abs(e - target)
            else:
                raise ValueError("In-correct Choose_Type")
            if ok:
                ewp_idx.append(i)
                dist.append(d)
    if not ewp_idx:  # (:308-315)
        ewp_idx = list(range(iteno))
        dist = [abs(sort_ritz[i] - target) for i in range(iteno)]
    jj = len(ewp_idx)
    ewno = min(jj, rppd)
    ritz_ev = np.zeros((iteno, rppd), complex)
    ritz_ew = np.zeros(rppd, complex)
    chosen = []
    plain = abs(target_org - target) < 1.0e-10 or choose_type_org == "CL"
    tor = float(np.real(target_org))
    for i in range(ewno):  # (:332-366)
        ind_sort = 0
        for k in range(1, jj):
            if dist[k] < dist[ind_sort]:
                if plain:
                    ind_sort = k
                elif choose_type_org == "RGTR":
                    e = ewl[ewp_idx[k]]
                    if e.real > tor and abs(e.imag) < 1.0e-12:
                        ind_sort = k
                elif choose_type_org == "RGTC":
                    if ewl[ewp_idx[k]].real > tor:
                        ind_sort = k
        if dist[ind_sort] > 1.0e12:
            ind_sort = min(range(jj), key=dist.__getitem__)
        chosen.append(ewp_idx[ind_sort])
        ritz_ew[i] = ew[ewp_idx[ind_sort]]
        v = VRR[:iteno, ewp_idx[ind_sort]]
        ritz_ev[:, i] = v / np.linalg.norm(v)
        dist[ind_sort] = 1.0e14
    if ewno < rppd:  # fill with the remaining Ritz pairs closest to target (:371-411)
        cand = [i for i in range(iteno) if abs(sort_ritz[i]) < 1.0e10 and i not in chosen]
        d2 = np.array([abs(ewl[i] - target) for i in cand])
        order = np.argsort(d2, kind="stable")
        add = min(rppd - ewno, len(order))
        for ii in range(add):
            i = cand[order[ii]]
            ritz_ew[ewno + ii] = ew[i]
            ritz_ev[:, ewno + ii] = VRR[:iteno, i] / np.linalg.norm(VRR[:iteno, i])
        ewno += add
    return ewno, ritz_ew[:max(ewno, 1)], ritz_ev[:, :max(ewno, 1)]


class _State:
    def __init__(self, ew_number, rppd):
        self.zflag_ew = np.zeros(ew_number + rppd, complex)
        self.flag_no_ew = 0
        self.dim_eigsp = 0
        self.total_inner_iter = 0
        self.log = []      # (ew_order, EPIteNo, ritz value, residual) per outer step (Locking4.m:184-195)
        self.locked = []   # (ew_order, EPIteNo, seconds, eigenvalue, residual) per lock (output.m)
        self.inner = []    # total inner iterations per locking call (Driver4.m:65)
        self.t0 = None


def _projected_eig(M, L):
    """eig(M, L) for Hermitian M > 0 and Hermitian L: L y = mu M y, lambda = 1/mu."""
    mu, Y = scipy.linalg.eigh(L, M, check_finite=False)
    with np.errstate(divide="ignore"):
        lam = np.where(mu != 0, 1.0 / mu, np.inf).astype(complex)
    Y = Y / np.linalg.norm(Y, axis=0, keepdims=True)  # unit 2-norm columns, as eig's QZ path returns
    return lam, Y.astype(complex)


def _locking(plan, A, B, EV, ew_order, init_vec, proj_dim, target, choose_type, rppd, no_restart, vec_r_tol,
             ew_number, coreq_org, ls, st, sigma_inner):
    """One call of GEP_AB_Herm_JDSIRA_Locking4 (Locking4.m:1-507)."""
    t = plan.torch
    L_len = init_vec.shape[1]
    max_iter_no = ls.get("maxit", 500)
    linear_system_tol = ls.get("tol", 1.0e-2)
    stop_flag = 0
    EPIteNo = 1
    iter_flag = 0
    pcg_iter_number = max_iter_no
    prev_rsdl = 1.0e2
    no_ew_near_tol = 0
    target_ritz_ew = target
    change_target = 0
    prev_ritz_ew = 1.0e14
    tol_ritz_ew = 5.0e-3
    choose_type_org = choose_type
    V = t.zeros((no_restart + 2, L_len), dtype=t.complex128, device=plan.device)
    Mm = np.zeros((no_restart + 2, no_restart + 2), complex)
    Lm = np.zeros((no_restart + 2, no_restart + 2), complex)
    V[:proj_dim] = init_vec[:proj_dim]
    ritz_ew = ritz_ev = ev_crnt = None
    ewno = 1
    iteno = proj_dim
    rsdl = 0.0
    while EPIteNo <= 6000 and stop_flag == 0:
        cols = []  # Rayleigh quotients (:91-103); columns stay on the device until all are done
        for ii in range(proj_dim):
            cols.append(plan.block_dot(V[ii:proj_dim], A(V[ii])))
            cols.append(plan.block_dot(V[ii:proj_dim], B(V[ii])))
        host = t.cat(cols).cpu().numpy() if cols else np.zeros(0, complex)
        off = 0
        for ii in range(proj_dim):
            w = proj_dim - ii
            Mm[ii:proj_dim, ii] = host[off:off + w]
            Lm[ii:proj_dim, ii] = host[off + w:off + 2 * w]
            off += 2 * w
            Mm[ii, ii + 1:proj_dim] = np.conj(Mm[ii + 1:proj_dim, ii])
            Lm[ii, ii + 1:proj_dim] = np.conj(Lm[ii + 1:proj_dim, ii])
        iteno = proj_dim
        while iteno <= no_restart and stop_flag == 0:
            ew, VRR = _projected_eig(Mm[:iteno, :iteno], Lm[:iteno, :iteno])
            if np.sum(np.isinf(ew)) > 1:
                raise RuntimeError("inf merg.")
            ewno, ritz_ew, ritz_ev = choose_ritz_pair(iteno, ew, ew_order, target_ritz_ew, VRR, vec_r_tol,
                                                      choose_type, rppd, target, choose_type_org, st.zflag_ew,
                                                      st.flag_no_ew)
            # u = V y / ||V y||, r = A u - theta B u, ||r|| (:170-178): one library call
            ev_crnt, rsdl_vec, rsdl = plan.jd_ritz_residual(V[:iteno], ritz_ev[:iteno, 0], ritz_ew[0])
            st.log.append((ew_order, EPIteNo, complex(ritz_ew[0]), rsdl))
            if 10.0 * vec_r_tol >= rsdl > vec_r_tol:
                no_ew_near_tol += 1
            else:
                no_ew_near_tol = 0
            if rsdl <= vec_r_tol or no_ew_near_tol >= 5:
                stop_flag = 1
            else:
                if rsdl < tol_ritz_ew and change_target == 1 and \
                        abs(prev_ritz_ew - ritz_ew[0]) / abs(ritz_ew[0]) < 5.0e-2:
                    target_ritz_ew = ritz_ew[0]
                    change_target = 0
                    choose_type = "CL"
                prev_ritz_ew = ritz_ew[0]
                if iteno + 1 <= st.dim_eigsp + no_restart:
                    if rsdl < 1.0e-1 and prev_rsdl / rsdl < 4.0:
                        pcg_iter_number = min(int(np.floor(max_iter_no * 1.2)), pcg_iter_number + 10)
                    prev_rsdl = rsdl
                    if ew_order == 1 and iteno <= 10 and rsdl > 1.0e-1:  # (:264-275)
                        pcg_iter_number = 500 if coreq_org.upper() == "JD" else 5000
                        iter_flag = 0
                    else:
                        if iter_flag == 0:
                            pcg_iter_number = max_iter_no
                            iter_flag = 1
                    # SIRA correction: inner MINRES on A (or A - sigma B), :311-312
                    sig = 0.0 if sigma_inner is None else float(sigma_inner)
                    # MINRES correction, rescaled when tiny (:318-320), orthogonalised against V by
                    # MGS in column order (:326-330) and normalised into V[iteno], then the new column
                    # of V^H A V and V^H B V (:335-341): one library call
                    its, mcol, lcol = plan.jd_expand(V, iteno, rsdl_vec, sig, linear_system_tol, pcg_iter_number)
                    st.total_inner_iter += int(its)
                    Mm[:iteno + 1, iteno], Lm[:iteno + 1, iteno] = mcol, lcol
                    Mm[iteno, :iteno] = np.conj(Mm[:iteno, iteno])
                    Lm[iteno, :iteno] = np.conj(Lm[:iteno, iteno])
            iteno += 1
            EPIteNo += 1
        if stop_flag == 0:  # restart (:361-383)
            Mm[:] = 0
            iteno -= 1
            jj = min(rppd - 1, ew_number + rppd - 1)
            cols = ritz_ev[:iteno, 1:jj + 1]
            for c in range(cols.shape[1]):
                init_vec[c] = plan.block_gemv(V[:iteno], t.as_tensor(cols[:, c]))
            jj = cols.shape[1]
            k = st.dim_eigsp
            V[k] = ev_crnt
            k += 1
            V[k:k + jj] = init_vec[:jj]
            k += jj
            proj_dim = gram_schmidt(plan, V[:k])
    # Ritz pairs beyond the first and their residuals (:389-416)
    tmp_rsdl = np.ones(max(ewno, 1) + 1)
    jj = 0
    for i in range(1, ewno):
        if jj + 1 <= ew_number + rppd - 1:
            v = plan.block_gemv(V[:iteno - 1], t.as_tensor(ritz_ev[:iteno - 1, i]))
            init_vec[jj] = v / _norm(plan, v)
            jj += 1
            rv = A(init_vec[jj - 1]) - ritz_ew[i] * B(init_vec[jj - 1])
            tmp_rsdl[i] = _norm(plan, rv)
    st.flag_no_ew += 1
    k = st.dim_eigsp
    u = ev_crnt / _norm_dev(plan, ev_crnt)  # lock the converged pair (:425-450)
    import time as _time
    st.locked.append((ew_order, EPIteNo, _time.perf_counter() - st.t0, complex(ritz_ew[0]), rsdl))
    V[k] = u
    EV[k] = u
    k += 1
    st.zflag_ew[st.flag_no_ew - 1] = ritz_ew[0]
    st.dim_eigsp += 1
    flag_cn = np.zeros(max(ewno, 1), int)
    j1 = j = 0
    for i in range(1, ewno):  # lock other converged Ritz pairs (:459-476)
        if i <= ew_number + rppd - 1 and tmp_rsdl[i] <= vec_r_tol and k < EV.shape[0]:
            V[k] = init_vec[i - 1]
            EV[k] = init_vec[i - 1]
            k += 1
            flag_cn[i] = 1
            j += 1
            st.flag_no_ew += 1
            st.zflag_ew[st.flag_no_ew - 1] = ritz_ew[i]
            st.dim_eigsp += 1
            j1 += 1
    j1 = 0
    for i in range(1, ewno):  # new search space (:483-492)
        if flag_cn[i] == 0 and j1 + 1 <= ew_number + rppd - 1:
            V[k] = init_vec[j1]
            k += 1
        j1 += 1
    i = gram_schmidt(plan, V[:k])
    proj_dim = min(i, ew_number + rppd - 1)
    init_vec[:proj_dim] = V[:proj_dim]
    return ew_order + j + 1, proj_dim, EPIteNo, stop_flag


def GEP_AB_Herm_JDSIRA_Driver4(plan, ew_number, stop_tolerance=1.0e-8, no_restart=50, target=0.0,
                               target_type="CL", CorEq="SIRA", LSinfo=None, initial_vec=None, sigma_inner=None,
                               seed=0):
    """JD-SIRA with locking for fncB x = omega fncA x (Driver4.m:1-81) on a plan with nb = 1.

    Returns a JDResult with the ``ew_number`` converged eigenvalues and eigenvectors."""
    if plan.nb != 1:
        raise ValueError("JD-SIRA runs one k-point per plan")
    t = plan.torch
    ls = {"solver": "minres", "tol": 1.0e-2, "maxit": 30000}
    if LSinfo:
        ls.update(LSinfo)
    L_len = 4 * plan.N
    A = plan.fncB
    B = plan.fncA
    if initial_vec is None:  # rand + i rand, normalised (main_EIG_np_sphere_n60_gamma3p46427.m:105-106)
        rng = np.random.default_rng(seed)
        v = rng.random(L_len) + 1j * rng.random(L_len)
        initial_vec = t.as_tensor(v / np.linalg.norm(v), device=plan.device).reshape(1, -1)
    rppd = max(4, initial_vec.shape[0])
    proj_dim = initial_vec.shape[0]
    init_vec = t.zeros((ew_number + rppd - 1, L_len), dtype=t.complex128, device=plan.device)
    init_vec[:proj_dim] = initial_vec
    EV = t.zeros((ew_number, L_len), dtype=t.complex128, device=plan.device)
    st = _State(ew_number, rppd)
    no_computed_ew = 1
    total_outer = 0
    rng = np.random.default_rng(seed + 1)
    import time as _time
    st.t0 = _time.perf_counter()
    with _one_blas_thread():
        no_computed_ew, total_outer = _driver_loop(plan, A, B, EV, no_computed_ew, init_vec, proj_dim, target,
                                                   target_type, rppd, no_restart, stop_tolerance, ew_number, CorEq,
                                                   ls, st, sigma_inner, rng, L_len, total_outer)
    res = JDResult(st.zflag_ew[:ew_number].copy(), EV, st.log, st.total_inner_iter, total_outer)
    res.locked = st.locked
    res.inner_per_lock = st.inner
    return res


def _driver_loop(plan, A, B, EV, no_computed_ew, init_vec, proj_dim, target, target_type, rppd, no_restart,
                 stop_tolerance, ew_number, CorEq, ls, st, sigma_inner, rng, L_len, total_outer):
    """The outer loop of Driver4.m:40-66: lock eigenpairs until ew_number have converged."""
    t = plan.torch
    while no_computed_ew <= ew_number:
        inner0 = st.total_inner_iter
        no_computed_ew, proj_dim, epit, stop_flag = _locking(plan, A, B, EV, no_computed_ew, init_vec, proj_dim, target,
                                                  target_type, rppd, no_restart, stop_tolerance, ew_number, CorEq,
                                                  ls, st, sigma_inner)
        total_outer += epit
        st.inner.append(st.total_inner_iter - inner0)
        if stop_flag == 0:
            raise NotConverged(FAME_ERR_NOT_CONVERGED, f"JD-SIRA: no convergence within 6000 outer steps "
                                                       f"({no_computed_ew - 1} of {ew_number} eigenpairs)")
        if proj_dim == 0:  # (Driver4.m:58-63)
            v = t.as_tensor(rng.standard_normal(L_len) + 0j, device=plan.device)
            init_vec[0] = v / np.sqrt(float(np.linalg.norm(v.cpu().numpy())))
            proj_dim = 1
    return no_computed_ew, total_outer

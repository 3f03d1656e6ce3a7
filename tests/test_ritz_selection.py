"""The vectorised Ritz-pair selection (jdsira.choose_ritz_pair) equals the loop-by-loop
restatement of Choose_Ritz_Pair_Unsymm4.m it replaced (kept below verbatim), on random
Ritz values with clusters at every threshold scale the reference tests (10^k tol for
k = -1 .. 5), converged eigenvalues among them, and all three selection types."""
import numpy as np
import pytest


def choose_ritz_pair_loops(iteno, ew, ew_order, target, VRR, vec_r_tol, choose_type, rppd, target_org, choose_type_org,
                     zflag_ew, flag_no_ew):
    """Select Ritz pairs (Choose_Ritz_Pair_Unsymm4.m:1-412); returns (ewno, ritz_ew, ritz_ev)."""
    ew = np.asarray(ew, complex)
    sort_ritz = ew.copy()
    sort_No = 5
    if ew_order > 1:
        sort_ew = np.zeros(iteno, complex)
        sort_ew[:flag_no_ew] = zflag_ew[:flag_no_ew]
        mult_cn_ew = np.zeros(flag_no_ew, int)
        for i in range(flag_no_ew):  # multiplicity of converged eigenvalues (:52-78)
            if sort_ew[i].real > -1.0e13:
                no_cn_ew = 1
                for i1 in range(i + 1, flag_no_ew):
                    if sort_ew[i1].real > -1.0e13:
                        k, flag_tol = -1, 1
                        while k <= 1 and flag_tol == 1:
                            if abs(sort_ew[i1] - sort_ew[i]) <= 10.0 ** k * vec_r_tol:
                                flag_tol = 0
                            else:
                                k += 1
                        if flag_tol == 0:
                            no_cn_ew += 1
                            sort_ew[i1] = -1.0e14
            else:
                no_cn_ew = 0
            mult_cn_ew[i] = no_cn_ew
        idx_ritz_w = np.zeros(iteno, int)  # 1-based links as in the reference (0 = first occurrence)
        mult_ritz_w = np.zeros(iteno, int)
        for i in range(iteno):  # multiplicity of Ritz values (:97-121)
            if idx_ritz_w[i] == 0:
                no_ritz_w = 1
                for i1 in range(i + 1, iteno):
                    if idx_ritz_w[i1] == 0:
                        k, flag_tol = -1, 1
                        while k <= 1 and flag_tol == 1:
                            if abs(sort_ritz[i1] - sort_ritz[i]) <= 10.0 ** k * vec_r_tol:
                                flag_tol = 0
                            else:
                                k += 1
                        if flag_tol == 0:
                            no_ritz_w += 1
                            idx_ritz_w[i1] = i + 1
            else:
                no_ritz_w = 0
            mult_ritz_w[i] = no_ritz_w
        for i in range(flag_no_ew):  # remove Ritz values equal to converged eigenvalues (:126-228)
            if sort_ew[i].real > -1.0e13:
                i1, flag, ind_sort = 0, 1, -1
                while i1 < iteno and flag == 1:
                    if idx_ritz_w[i1] == 0:
                        k, flag_tol = -1, 1
                        while k <= sort_No and flag_tol == 1:
                            if abs(sort_ritz[i1] - sort_ew[i]) <= 10.0 ** k * vec_r_tol:
                                flag_tol = 0
                                ind_sort = i1
                            else:
                                k += 1
                        if flag_tol == 0:
                            flag = 0
                        else:
                            i1 += 1
                    else:
                        i1 += 1
                if flag == 1:
                    continue  # 'Error in Choose_Ritz_Pair' (printed by the reference, no action)
                if mult_cn_ew[i] == mult_ritz_w[ind_sort]:
                    sort_ritz[ind_sort] = 1.0e14
                    sort_ritz[idx_ritz_w == ind_sort + 1] = 1.0e14
                    idx_ritz_w[ind_sort] = -1
                elif mult_cn_ew[i] < mult_ritz_w[ind_sort]:
                    idx_mult = [ind_sort]
                    ld = [np.linalg.norm(VRR[ew_order - 1:iteno, ind_sort])]
                    for jj in range(iteno):
                        if idx_ritz_w[jj] == ind_sort + 1:
                            idx_mult.append(jj)
                            ld.append(np.linalg.norm(VRR[ew_order - 1:iteno, jj]))
                    ld = np.array(ld)
                    for _ in range(mult_cn_ew[i]):
                        kmin = 0
                        tmp = ld[0]
                        for i2 in range(1, mult_ritz_w[ind_sort]):
                            if tmp > ld[i2]:
                                tmp, kmin = ld[i2], i2
                        sort_ritz[idx_mult[kmin]] = 1.0e14
                        ld[kmin] = 1.0e14
                    idx_ritz_w[ind_sort] = -1
                else:
                    sort_ritz[ind_sort] = 1.0e14
                    sort_ritz[idx_ritz_w == ind_sort + 1] = 1.0e14
                    idx_ritz_w[ind_sort] = -1
    # candidates (:235-306)
    ewp_idx, dist = [], []
    for i in range(iteno):
        if abs(sort_ritz[i]) < 1.0e10:
            if choose_type == "CL":
                ok = (abs(ew[i].real) < 1e-4 and ew[i].imag > 1e-3) or (ew[i].real > np.real(target)
                                                                          and abs(ew[i].imag) < 1e-8)
                d = abs(ew[i].real - np.real(target))
            elif choose_type == "RGTR":
                ok = ew[i].real > np.real(target) and abs(ew[i].imag) < 1.0e-12
                d = abs(ew[i].real - np.real(target))
            elif choose_type == "RGTC":
                ok = ew[i].real > np.real(target)
                d = abs(ew[i] - target)
            else:
                raise ValueError("In-correct Choose_Type")
            if ok:
                ewp_idx.append(i)
                dist.append(d)
    if not ewp_idx:  # (:308-315)
        ewp_idx = list(range(iteno))
        dist = [abs(sort_ritz[i] - target) for i in range(iteno)]
    jj = len(ewp_idx)
    dist = np.array(dist, float)
    ewno = min(jj, rppd)
    ritz_ev = np.zeros((iteno, rppd), complex)
    ritz_ew = np.zeros(rppd, complex)
    chosen = []
    for i in range(ewno):  # (:332-366)
        ind_sort = 0
        for k in range(1, jj):
            if dist[k] < dist[ind_sort]:
                if abs(target_org - target) < 1.0e-10 or choose_type_org == "CL":
                    ind_sort = k
                elif choose_type_org == "RGTR":
                    if ew[ewp_idx[k]].real > np.real(target_org) and abs(ew[ewp_idx[k]].imag) < 1.0e-12:
                        ind_sort = k
                elif choose_type_org == "RGTC":
                    if ew[ewp_idx[k]].real > np.real(target_org):
                        ind_sort = k
        if dist[ind_sort] > 1.0e12:
            ind_sort = int(np.argmin(dist[:jj]))
        chosen.append(ewp_idx[ind_sort])
        ritz_ew[i] = ew[ewp_idx[ind_sort]]
        v = VRR[:iteno, ewp_idx[ind_sort]]
        ritz_ev[:, i] = v / np.linalg.norm(v)
        dist[ind_sort] = 1.0e14
    if ewno < rppd:  # fill with the remaining Ritz pairs closest to target (:371-411)
        out_idx = [i for i in range(iteno) if i not in chosen]
        cand = [i for i in range(iteno) if abs(sort_ritz[i]) < 1.0e10 and i in out_idx]
        d2 = np.array([abs(ew[i] - target) for i in cand])
        order = np.argsort(d2, kind="stable")
        add = min(rppd - ewno, len(order))
        for ii in range(add):
            i = cand[order[ii]]
            ritz_ew[ewno + ii] = ew[i]
            ritz_ev[:, ewno + ii] = VRR[:iteno, i] / np.linalg.norm(VRR[:iteno, i])
        ewno += add
    return ewno, ritz_ew[:max(ewno, 1)], ritz_ev[:, :max(ewno, 1)]


def _case(rng, iteno, nconv, tol):
    base = rng.uniform(0.05, 2.0, iteno)
    ew = base.astype(complex)
    # clusters: copy a value with offsets at each threshold scale (and slightly above it)
    for _ in range(iteno // 3):
        a, b = rng.integers(0, iteno, 2)
        k = rng.integers(-2, 7)
        ew[b] = ew[a] + (10.0 ** k) * tol * rng.choice([0.5, 0.999, 1.001, 2.0]) * np.exp(1j * rng.uniform(0, 6.3))
    ew[rng.integers(0, iteno, 2)] += 1j * rng.uniform(-1e-3, 1e-2, 2)        # a few complex Ritz values
    conv = np.zeros(nconv + 5, complex)
    pick = rng.choice(iteno, nconv, replace=False)
    conv[:nconv] = ew[pick] + (10.0 ** rng.integers(-2, 6, nconv)) * tol * 0.7
    if nconv > 2:
        conv[1] = conv[0]                                                        # repeated converged value
    VRR = rng.standard_normal((iteno, iteno)) + 1j * rng.standard_normal((iteno, iteno))
    return ew, VRR, conv


def _cluster_case(rng, tol):
    """Ritz clusters at the target with members spaced at 10^k tol scales, and converged
    eigenvalues sitting on cluster members, so that multiplicities decide what is removed
    and hence which pairs are selected."""
    target = 0.3
    c = target + rng.uniform(1e-3, 5e-3)
    scales = 10.0 ** rng.integers(-2, 7, 3) * rng.choice([0.3, 0.9, 3.0], 3)
    cluster = c + np.concatenate([[0.0], np.cumsum(scales)]) * tol
    others = rng.uniform(0.5, 2.0, int(rng.integers(2, 12)))
    ew = np.concatenate([cluster, others]).astype(complex)
    ew = ew[rng.permutation(ew.size)]
    nconv = int(rng.integers(1, 4))
    conv = np.zeros(nconv + 5, complex)
    conv[:nconv] = cluster[rng.integers(0, cluster.size, nconv)] + rng.choice([0.0, 0.5, 5.0]) * tol
    iteno = ew.size
    VRR = rng.standard_normal((iteno, iteno)) + 1j * rng.standard_normal((iteno, iteno))
    return ew, VRR, conv, nconv, target


@pytest.mark.parametrize("seed", range(200))
def test_clustered_selection_equals_loops(pkg, seed):
    rng = np.random.default_rng(1000 + seed)
    tol = 10.0 ** rng.uniform(-10, -7)
    ew, VRR, conv, nconv, target = _cluster_case(rng, tol)
    iteno = ew.size
    for ctype in ("CL", "RGTR", "RGTC"):
        args = (iteno, ew, nconv + 1, target, VRR, tol, ctype, 3, target, ctype, conv, nconv)
        a = pkg.jdsira.choose_ritz_pair(*args)
        b = choose_ritz_pair_loops(*args)
        assert a[0] == b[0]
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[2], b[2])


# chained clusters (offsets in units of vec_r_tol, in array order): a later root that is
# close to a member already linked to an earlier root must not take it over
CHAINS = [([0, 16, 8], [16]), ([0, 16, 8], [0]), ([0, 16, 8, 24], [16, 16.5]), ([5, 0, 13, 9], [13]),
          ([0, 9, 18, 27], [9, 18]), ([0, 30, 8, 22, 15], [30]), ([0, 0.05, 0.5, 7], [0.05, 7])]


@pytest.mark.parametrize("offsets,conv_off", CHAINS)
@pytest.mark.parametrize("tol", [1e-9, 1e-7])
def test_chained_clusters_equal_loops(pkg, offsets, conv_off, tol):
    rng = np.random.default_rng(len(offsets))
    c, target = 0.302, 0.3
    ew = np.concatenate([c + np.array(offsets, float) * tol, [0.9, 1.4]]).astype(complex)
    conv = np.zeros(len(conv_off) + 5, complex)
    conv[:len(conv_off)] = c + np.array(conv_off, float) * tol
    iteno = ew.size
    VRR = rng.standard_normal((iteno, iteno)) + 1j * rng.standard_normal((iteno, iteno))
    for ctype in ("CL", "RGTR", "RGTC"):
        args = (iteno, ew, len(conv_off) + 1, target, VRR, tol, ctype, 3, target, ctype, conv, len(conv_off))
        try:
            b = choose_ritz_pair_loops(*args)
        except IndexError:   # the reference's own indexing fails on this layout; so must ours
            with pytest.raises(IndexError):
                pkg.jdsira.choose_ritz_pair(*args)
            continue
        a = pkg.jdsira.choose_ritz_pair(*args)
        assert a[0] == b[0]
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("seed", range(40))
def test_vectorised_selection_equals_loops(pkg, seed):
    rng = np.random.default_rng(seed)
    tol = 10.0 ** rng.uniform(-10, -6)
    iteno = int(rng.integers(3, 52))
    nconv = int(rng.integers(0, min(iteno, 8)))
    ew, VRR, conv = _case(rng, iteno, nconv, tol)
    ew_order = nconv + 1
    for ctype in ("CL", "RGTR", "RGTC"):
        for target in (0.3, 1.0):
            args = (iteno, ew, ew_order, target, VRR, tol, ctype, 3, target, ctype, conv, nconv)
            a = pkg.jdsira.choose_ritz_pair(*args)
            b = choose_ritz_pair_loops(*args)
            assert a[0] == b[0]
            assert np.array_equal(a[1], b[1])
            assert np.array_equal(a[2], b[2])

"""Pins of the CPU oracle against what the mathematics fixes (no reference run exists).

Each test is chosen so that a plausible slip in the restatement (a dropped
term, a sign, a conjugate, a transposed operand, a wrong wrap phase) fails it.
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

rng = np.random.default_rng(7)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def herm_tensors(n):
    X = crand(n, 3, 3)
    return X + np.conj(np.transpose(X, (0, 2, 1))) + 6 * np.eye(3)[None]


@pytest.mark.parametrize("m", [1, 3, 4, 7])
def test_K_QPBC_svd_closed_form(fo, m):
    """K_QPBC = V diag(lambda) V^H with V = diag(S) F  (SVD_Matrix_1D.m:49-59, :91-93)."""
    theta, delta = 0.7, 1.0 / m
    K = fo.K_QPBC(m, theta, delta).toarray()
    V = np.diag(fo.S_sDFT(m, theta)) @ fo.DFT_mtx(m)
    lam = fo.lambda_quasiperiodic(m, theta, delta)
    assert np.allclose(V.conj().T @ V, np.eye(m), atol=1e-13)
    assert np.linalg.norm(K - V @ np.diag(lam) @ V.conj().T) <= 1e-12 * np.linalg.norm(K)


def test_MtxVecMul_FFT_matches_dense_dft(fo):
    """Psi = diag(S) kron(F_z, F_y, F_x) per component; Psis is its adjoint."""
    dims = (3, 4, 5)
    theta = np.array([0.4, -1.3, 2.2])
    A = fo.svd_matrix_generate(dims, theta, 1.0 / np.array(dims))
    F = np.kron(fo.DFT_mtx(5), np.kron(fo.DFT_mtx(4), fo.DFT_mtx(3)))
    S1 = A["S"][: np.prod(dims)]
    Psi1 = np.diag(S1) @ F
    x = crand(3 * np.prod(dims))
    y = fo.MtxVecMul_FFT(x, dims, "notranspose", A["S"])
    N = np.prod(dims)
    for c in range(3):
        assert np.allclose(y[c * N:(c + 1) * N], Psi1 @ x[c * N:(c + 1) * N], atol=1e-13)
    z = fo.MtxVecMul_FFT(x, dims, "transpose", A["S"])
    for c in range(3):
        assert np.allclose(z[c * N:(c + 1) * N], Psi1.conj().T @ x[c * N:(c + 1) * N], atol=1e-13)


@pytest.mark.parametrize("dims", [(3, 4, 5), (6, 6, 6), (4, 2, 7)])
def test_curl_factorisation(fo, dims):
    """C = Pr diag(Sigma_r) Qr^H and C^H = Qr diag(Sigma_r) Pr^H with the explicit curl
    (Funhand_Generate_Lanzcos.m:218-219), plus Qrs Qr = Prs Pr = I (:213-214)."""
    theta = np.array([0.3, -1.1, 2.0])
    delta = 1.0 / np.array(dims)
    A = fo.svd_matrix_generate(dims, theta, delta)
    f = fo.Fnc(A)
    C = fo.curl_C(dims, delta, theta)
    N = np.prod(dims)
    x = crand(3 * N)
    cn = sp.linalg.norm(C)
    assert np.linalg.norm(C @ x - f.Pr(A["Sigma_r"] * f.Qrs(x))) <= 1e-13 * cn * np.linalg.norm(x)
    assert np.linalg.norm(C.conj().T @ x - f.Qr(A["Sigma_r"] * f.Prs(x))) <= 1e-13 * cn * np.linalg.norm(x)
    y = crand(2 * N)
    assert np.linalg.norm(f.Qrs(f.Qr(y)) - y) <= 1e-13 * np.linalg.norm(y)
    assert np.linalg.norm(f.Prs(f.Pr(y)) - y) <= 1e-13 * np.linalg.norm(y)


def test_Pi_closed_form_cross_products(fo):
    """Pi_1 = conj(l) x w / |conj(l) x w|, Pi_2 = -(conj(l) x conj(Pi_1))/|l| (SVD_Matrix_curl.m:55-57)
    and Lambda_r Pi_0 = 0 (Pi_0 spans the null space)."""
    dims = (4, 3, 5)
    theta = np.array([1.0, 0.2, -0.4])
    A = fo.svd_matrix_generate(dims, theta, 1.0 / np.array(dims))
    lam = np.stack([A["Lx"], A["Ly"], A["Lz"]], 1)
    w = np.array([1, 2, 3]) / np.sqrt(14)
    c = np.conj(lam)
    p1 = np.cross(c, w)
    p1 /= np.linalg.norm(p1, axis=1)[:, None]
    p2 = -np.cross(c, np.conj(p1)) / np.linalg.norm(lam, axis=1)[:, None]
    N = np.prod(dims)
    assert np.allclose(A["D_Pi_1"], p1.T.reshape(-1), atol=1e-14)
    assert np.allclose(A["D_Pi_2"], p2.T.reshape(-1), atol=1e-14)
    assert np.linalg.norm(A["Lambda_r"] @ A["Pi_0"].toarray()) < 1e-10
    assert A["Qr"].shape == (3 * N, 2 * N)


def test_material_nnz_and_vacuum(fo):
    """nnz/row 9/25/20; vacuum (identity tensors) gives N_int = M_int = I, J_int = 0."""
    dims = (4, 5, 3)
    N = np.prod(dims)
    theta = np.array([0.3, 0.9, -2.0])
    T = herm_tensors(N)
    B = fo.interpolate_material(dims, theta, T, herm_tensors(N), crand(N, 3, 3))
    for key, nnz in (("N_int", 9), ("M_int", 25), ("J_int", 20)):
        c = np.diff(B[key].tocsr().indptr)
        assert c.min() == nnz and c.max() == nnz
    I = np.broadcast_to(np.eye(3), (N, 3, 3)).astype(complex)
    B0 = fo.interpolate_material(dims, theta, I, I, 0 * I)
    assert abs(B0["N_int"] - sp.identity(3 * N)).max() < 1e-15
    assert abs(B0["M_int"] - sp.identity(3 * N)).max() < 1e-15
    assert abs(B0["J_int"]).max() == 0


@pytest.mark.parametrize("hermitian", [True, False])
def test_factored_stencil_equals_sparse(fo, hermitian):
    """The vertex gather / 6x6 / scatter form (what the device evaluates) equals
    [M h + J^H e; J h + N e] with the reference's sparse interpolated matrices."""
    dims = (5, 4, 6)
    N = np.prod(dims)
    theta = np.array([0.7, -2.5, 1.3])
    if hermitian:
        Nv, Mv = herm_tensors(N), herm_tensors(N)
    else:
        Nv, Mv = crand(N, 3, 3), crand(N, 3, 3)
    Jv = crand(N, 3, 3)
    B = fo.interpolate_material(dims, theta, Nv, Mv, Jv)
    h, e = crand(3 * N), crand(3 * N)
    yh, ye = fo.material_apply_factored(dims, theta, Nv, Mv, Jv, h, e)
    yh_r = B["M_int"] @ h + B["J_int"].conj().T @ e
    ye_r = B["J_int"] @ h + B["N_int"] @ e
    assert np.linalg.norm(yh - yh_r) <= 1e-14 * np.linalg.norm(yh_r)
    assert np.linalg.norm(ye - ye_r) <= 1e-14 * np.linalg.norm(ye_r)


def test_fncB_vacuum_is_identity_and_hermitian_pd(fo):
    dims = (4, 4, 4)
    N = np.prod(dims)
    theta = np.array([0.5, 0.3, 0.1]) * np.pi
    I = np.broadcast_to(np.eye(3), (N, 3, 3)).astype(complex)
    P0 = fo.Problem(dims, theta, I, I, 0 * I)
    x = crand(4 * N)
    assert np.linalg.norm(P0.fncB(x) - x) <= 1e-13 * np.linalg.norm(x)
    T = herm_tensors(N)
    P = fo.Problem(dims, theta, T, herm_tensors(N), 0.3 * crand(N, 3, 3))
    y = crand(4 * N)
    a, b = np.vdot(x, P.fncB(y)), np.conj(np.vdot(y, P.fncB(x)))
    assert abs(a - b) <= 1e-12 * abs(a)


def _woodpile_problem(fo, dims, gamma=0.2, k=(0.5, 0.3, 0.1)):
    import __graft_entry__ as ge

    S = ge.load_package().structures
    theta = S.theta_of_k(np.pi * np.asarray(k), S.SC_LATTICE)
    table = S.bi_isotropic_table(13.0, gamma, S.SC_LATTICE)
    ids = S.woodpile_sc_ids(dims)
    T = table[ids]
    return fo.Problem(dims, theta, T[:, 0], T[:, 1], T[:, 2]), theta


def _dense(op, n):
    return np.stack([op(e) for e in np.eye(n, dtype=complex)], 1)


def test_nfsep_spectrum_equals_original_curl_problem(fo):
    """NFSEP  fncB x = w fncA x  (SVD machinery + fncA + fncB) has exactly the non-zero
    spectrum of the original problem  B_mat [C^H h; -C e] = i w [e; h]  built only from
    the explicit curl (EvaluteError_Chiral_no_inv.m:43-44)."""
    dims = (3, 3, 4)
    P, theta = _woodpile_problem(fo, dims, gamma=0.5)
    N = P.N
    Ad = _dense(P.fncA, 4 * N)
    Bd = _dense(P.fncB, 4 * N)
    d = scipy.linalg.eigh(Ad, Bd, eigvals_only=True)
    w_nf = np.sort(1.0 / d)
    C = fo.curl_C(dims, P.delta, theta).toarray()
    Bm = P.mtx_B
    Nn, J, M = Bm["N_int"].toarray(), Bm["J_int"].toarray(), Bm["M_int"].toarray()
    Bmat = np.block([[Nn, J], [J.conj().T, M]])
    Z = np.zeros((3 * N, 3 * N))
    T = -1j * Bmat @ np.block([[Z, C.conj().T], [-C, Z]])
    w_org = np.linalg.eigvals(T)
    w_org = w_org[np.abs(w_org) > 1e-6 * np.abs(w_org).max()]
    assert len(w_org) == 4 * N
    assert np.max(np.abs(w_org.imag)) < 1e-8 * np.abs(w_org).max()
    w_org = np.sort(w_org.real)
    assert np.allclose(w_nf, w_org, rtol=1e-9, atol=1e-9 * np.abs(w_org).max())


def test_lanczos_vacuum_closed_form(fo):
    """Vacuum: omega = 1/d are the smallest Sigma_r, i.e. sqrt(sum_d |(e^{i(2 pi j+theta)/n}-1)/delta|^2).

    Sigma_r = [sqrt(Lq); sqrt(Lq)] makes every value at least doubly degenerate; a
    single-vector Krylov method sees each eigenspace once, so the k smallest
    DISTINCT values are expected."""
    dims = (4, 4, 4)
    N = np.prod(dims)
    theta = np.array([0.5, 0.3, 0.1]) * np.pi
    I = np.broadcast_to(np.eye(3), (N, 3, 3)).astype(complex)
    P = fo.Problem(dims, theta, I, I, 0 * I)
    k = 4
    v0 = np.ones(4 * N, complex)
    ev, ew, _ = fo.lanczos(P.fncA, P.fncB, lambda r: P.fncinvB(r)[0], 4 * N, k, 1e-10, 100, 2 * k, v0)
    omega = np.sort(1.0 / ew.real)
    lam = [np.abs((np.exp(1j * (2 * np.pi * np.arange(n) + theta[d]) / n) - 1) * n) ** 2 for d, n in enumerate(dims)]
    sig = np.sqrt(lam[0][None, None, :] + lam[1][None, :, None] + lam[2][:, None, None]).ravel()
    u = np.sort(sig)
    distinct = u[np.concatenate([[True], np.diff(u) > 1e-9 * u[-1]])]
    assert np.allclose(omega, distinct[:k], rtol=1e-9)


def test_lanczos_matches_dense_gep(fo):
    """Woodpile at n=4: Lanczos eigenvalues equal dense eigh of the GEP; residual small."""
    dims = (4, 4, 4)
    P, _ = _woodpile_problem(fo, dims)
    N = P.N
    k = 4
    v0 = np.ones(4 * N, complex)
    ev, ew, _ = fo.lanczos(P.fncA, P.fncB, lambda r: P.fncinvB(r)[0], 4 * N, k, 1e-10, 200, 2 * k, v0)
    d = scipy.linalg.eigh(_dense(P.fncA, 4 * N), _dense(P.fncB, 4 * N), eigvals_only=True)
    expect = np.sort(d)[::-1][:k]
    assert np.allclose(np.sort(ew.real)[::-1], expect, rtol=1e-10)
    for j in range(k):
        x = ev[:, j]
        r = P.fncA(x) - ew[j] * P.fncB(x)
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(x) * np.abs(ew[j])


def test_restoration_satisfies_original_problem(fo):
    dims = (4, 4, 4)
    P, theta = _woodpile_problem(fo, dims)
    N = P.N
    k = 2
    v0 = np.ones(4 * N, complex)
    ev, ew, _ = fo.lanczos(P.fncA, P.fncB, lambda r: P.fncinvB(r)[0], 4 * N, k, 1e-10, 200, 2 * k, v0)
    omega = 1.0 / ew
    e, h = fo.eigen_restoration(ev, P.fnc, P.mtx_B)
    C = fo.curl_C(dims, P.delta, theta)
    r_nf, r_org = fo.evaluate_error(ev, omega, e, h, C, P.fncA, P.fncB, P.mtx_B)
    for j in range(k):
        assert r_org[j] <= 1e-7 * np.abs(omega[j]) * np.linalg.norm(np.concatenate([e[:, j], h[:, j]]))


def test_pcg_residual(fo):
    dims = (4, 3, 5)
    P, _ = _woodpile_problem(fo, dims)
    x = crand(4 * P.N)
    y, flag, relres, it = fo.pcg(P.fncB, x, 1e-12, 1000)
    assert flag == 0
    assert np.linalg.norm(P.fncB(y) - x) <= 1e-12 * np.linalg.norm(x) * 1.01


@pytest.mark.parametrize("shifted", [False, True])
def test_minres_iterates_are_krylov_least_squares(fo, shifted):
    """x_k = argmin ||b - A x|| over K_k(A, b): brute force with an explicit basis.
    Shifted: A Hermitian indefinite (the A - sigma B inner operator)."""
    rng = np.random.default_rng(11)
    n = 40
    Q = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    d = rng.uniform(0.5, 3.0, n) * (np.where(np.arange(n) % 3 == 0, -1.0, 1.0) if shifted else 1.0)
    A = Q @ np.diag(d) @ Q.conj().T
    A = 0.5 * (A + A.conj().T)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for k in (1, 2, 5, 9):
        x, flag, relres, it = fo.minres(lambda v: A @ v, b, 1e-30, k)
        assert it == k and flag == 1
        K = np.empty((n, k), complex)
        K[:, 0] = b
        for j in range(1, k):
            K[:, j] = A @ K[:, j - 1]
        Kq = np.linalg.qr(K)[0]
        y = np.linalg.lstsq(A @ Kq, b, rcond=None)[0]
        xls = Kq @ y
        assert np.linalg.norm(x - xls) <= 1e-9 * np.linalg.norm(xls)
        # the recursive residual estimate is the true residual
        assert abs(relres - np.linalg.norm(b - A @ x) / np.linalg.norm(b)) < 1e-9
    x, flag, relres, it = fo.minres(lambda v: A @ v, b, 1e-12, 500)
    assert flag == 0 and np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-10 * np.linalg.norm(x)


def test_vertex_material_closed_form(fo):
    """Per-vertex slots (general batched inverse) = the bi-isotropic closed form n = mu/d,
    m = eps/d, j = -i gamma/d (d = eps mu - gamma^2) times A'A; derivatives likewise."""
    rng = np.random.default_rng(8)
    e = rng.uniform(1, 13, 7)
    g = rng.uniform(0, 0.9, 7)
    mu = 1.3
    A = rng.standard_normal((3, 3))
    G = A.T @ A
    N, M, J = fo.vertex_material(e, g, mu, A)
    d = e * mu - g ** 2
    assert np.allclose(N, (mu / d)[:, None, None] * G, rtol=1e-13, atol=0)
    assert np.allclose(M, (e / d)[:, None, None] * G, rtol=1e-13, atol=0)
    assert np.allclose(J, (-1j * g / d)[:, None, None] * G, rtol=1e-13, atol=0)
    h = 1e-6
    for wrt, (ep, gp) in (("eps", (e + h, g)), ("gamma", (e, g + h))):
        em, gm = (e - h, g) if wrt == "eps" else (e, g - h)
        dN, dM, dJ = fo.vertex_material_derivative(e, g, mu, A, wrt)
        Np, Mp, Jp = fo.vertex_material(ep, gp, mu, A)
        Nm, Mm, Jm = fo.vertex_material(em, gm, mu, A)
        for D, Pp, Pm in ((dN, Np, Nm), (dM, Mp, Mm), (dJ, Jp, Jm)):
            assert np.allclose(D, (Pp - Pm) / (2 * h), rtol=1e-7, atol=1e-9)

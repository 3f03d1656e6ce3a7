"""Pins of the FCC skewed-grid oracle (oracle/fcc_oracle.py) by what the mathematics
fixes, independent of the restated reshapes:

* T is unitary and TH is its adjoint;
* T equals the closed-form triple sum of twisted inverse DFTs (einsum with explicit
  DFT matrices, no reshapes);
* the explicit sparse difference matrices of Construct_GEP_mtx_A.m are diagonalised
  by T with the Lambda's of SVDSingleCurl_FCC.m (two reference files agreeing);
* each curl block factors as C_v (I_3 (x) T) Q_v = (I_3 (x) T) P_v diag(Sigma_v), with
  Q_v, P_v orthonormal;
* the parameter matrices are four permuted copies of pointwise 3x3 tensors.
"""
import numpy as np
import pytest

from oracle import fcc_oracle as fc

SIZES = [(2, 3, 3), (4, 6, 6), (6, 3, 4)]
KS = [np.array([0.31, -0.17, 0.23]), np.array([1.3, 0.4, -0.9])]


def setup(dims, k):
    n1, n2, n3 = dims
    a1, a2, a3 = fc.fcc_lattice()
    delta = np.array([1 / n1, np.sqrt(3) / 2 / n2, np.sqrt(6) / 3 / n3])
    return fc.fft_tables(n1, n2, n3, k, a1, a2, a3, delta), (a1, a2, a3), delta


def dense(f, n):
    return np.stack([f(e) for e in np.eye(n, dtype=complex)], 1)


@pytest.mark.parametrize("dims", SIZES)
def test_T_unitary(dims):
    tab, _, _ = setup(dims, KS[0])
    n = int(np.prod(dims))
    Tm = dense(lambda v: fc.T(v, tab), n)
    THm = dense(lambda v: fc.TH(v, tab), n)
    assert np.abs(Tm.conj().T @ Tm - np.eye(n)).max() < 1e-13
    assert np.abs(THm - Tm.conj().T).max() < 1e-13


@pytest.mark.parametrize("dims", SIZES)
def test_T_closed_form(dims):
    n1, n2, n3 = dims
    tab, _, _ = setup(dims, KS[1])
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n1 * n2 * n3) + 1j * rng.standard_normal(n1 * n2 * n3)
    X = x.reshape(n1, n2, n3)           # X[m, nn, l]: z-fastest input
    W = lambda n: np.exp(2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n) / n  # noqa: E731
    # z: sum_l e^{2 pi i j3 l/n3} X[m,nn,l], then D_jell[j3,nn,m]
    Z = np.einsum("jl,mnl->mnj", W(n3), X) * np.transpose(tab["D_jell"], (2, 1, 0))
    # y: sum_nn e^{2 pi i j2 nn/n2}, then D_jx[j2,m]
    Y = np.einsum("kn,mnj->mkj", W(n2), Z) * tab["D_jx"].T[:, :, None]
    # x: sum_m e^{2 pi i j1 m/n1}, then D_kx[j1]
    Xo = np.einsum("im,mkj->ikj", W(n1), Y) * tab["D_kx"][:, None, None]
    want = np.sqrt(n1 * n2 * n3) * np.transpose(Xo, (2, 1, 0)).ravel()   # x-fastest output
    assert np.abs(fc.T(x, tab) - want).max() < 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("dims", SIZES)
@pytest.mark.parametrize("ki", [0, 1])
def test_difference_matrices_diagonalised_by_T(dims, ki):
    tab, (a1, a2, a3), delta = setup(dims, KS[ki])
    n = int(np.prod(dims))
    Tm = dense(lambda v: fc.T(v, tab), n)
    _, G = fc.curl_blocks(*dims, KS[ki], a1, a2, a3, delta)
    for Gd, L in zip(G, (tab["Lambda_x"], tab["Lambda_y"], tab["Lambda_z"])):
        assert np.abs(Gd @ Tm - Tm * L[None, :]).max() < 1e-12 * np.abs(L).max()


@pytest.mark.parametrize("dims", SIZES[:2])
def test_curl_blocks_factor(dims):
    k = KS[0]
    tab, (a1, a2, a3), delta = setup(dims, k)
    n = int(np.prod(dims))
    Tm = dense(lambda v: fc.T(v, tab), n)
    I3T = np.kron(np.eye(3), Tm)
    blocks, _ = fc.curl_blocks(*dims, k, a1, a2, a3, delta)
    svd = fc.svd_curl(tab)
    for v, C in enumerate(blocks):
        Q1, Q2 = svd["pairs"][v]
        Q = np.hstack([fc._pointwise(Q1).toarray(), fc._pointwise(Q2).toarray()])
        P = np.hstack([fc._pointwise(-np.conj(Q2)).toarray(), fc._pointwise(np.conj(Q1)).toarray()])
        assert np.abs(Q.conj().T @ Q - np.eye(2 * n)).max() < 1e-13
        assert np.abs(P.conj().T @ P - np.eye(2 * n)).max() < 1e-13
        sig = svd["Sigma_r"][2 * v * n:(2 * v + 2) * n]
        assert np.abs(C @ I3T @ Q - (I3T @ P) * sig[None, :]).max() < 1e-12 * sig.max()
    # Lambda_Qr / Lambda_Pr are the block diagonals [LQ, LP] and [LP, LQ]
    LQr, LPr = svd["Lambda_Qr"].toarray(), svd["Lambda_Pr"].toarray()
    assert LQr.shape == (24 * n, 16 * n)
    assert np.array_equal(LQr[:12 * n, :8 * n], LPr[12 * n:, 8 * n:])
    assert np.array_equal(LQr[12 * n:, 8 * n:], LPr[:12 * n, :8 * n])
    assert not LQr[:12 * n, 8 * n:].any() and not LPr[12 * n:, :8 * n].any()


def test_parameter_matrix_is_four_permuted_pointwise_copies():
    rng = np.random.default_rng(9)
    n = 30
    sets = [rng.random(n) < 0.4 for _ in range(4)]
    Tin = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    Tair = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    M = fc.parameter_mtx(n, sets, Tin, Tair).toarray()
    assert np.count_nonzero(M) == 36 * n
    # group the 12 n-blocks into copies: blocks coupled by a nonzero entry
    blk = (np.abs(M.reshape(12, n, 12, n)).sum(axis=(1, 3)) > 0)
    copies, seen = [], set()
    for b in range(12):
        if b in seen:
            continue
        c = sorted(np.nonzero(blk[b])[0])
        assert len(c) == 3 and all(sorted(np.nonzero(blk[j])[0]) == c for j in c)
        copies.append(c)
        seen |= set(c)
    assert sorted(sum(copies, [])) == list(range(12))
    # per point: each copy's 3x3 block, in tensor order, is T_in or T_air (one set per copy)
    n_in = 0
    for c in copies:
        for p in range(n):
            B = M[np.ix_([b * n + p for b in c], [b * n + p for b in c])]
            # find the component order that makes B equal one of the tensors
            import itertools
            hits = [(perm, which) for perm in itertools.permutations(range(3)) for which, Tt in (("in", Tin), ("air", Tair))
                    if np.array_equal(B[np.ix_(perm, perm)], Tt)]
            assert hits, (c, p)
            n_in += hits[0][1] == "in"
    assert 0 < n_in < 4 * n


def test_gep_B_blocks():
    rng = np.random.default_rng(4)
    n = 12
    es = [rng.random(n) < 0.5 for _ in range(4)]
    fs = [rng.random(n) < 0.5 for _ in range(4)]
    ten = [rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)) for _ in range(8)]
    m = fc.eps_mu_zeta_xi(n, es, fs, ten[:4], ten[4:])
    B = fc.gep_B(m).toarray()
    blk = lambda i, j: B[12 * n * i:12 * n * (i + 1), 12 * n * j:12 * n * (j + 1)]  # noqa: E731
    assert np.array_equal(blk(0, 1), m["Zeta_f"].toarray()) and np.array_equal(blk(0, 2), m["Mu_f"].toarray())
    assert np.array_equal(blk(1, 0), m["Zeta_e"].toarray()) and np.array_equal(blk(1, 3), m["Mu_e"].toarray())
    assert np.array_equal(blk(2, 0), -m["Eps_e"].toarray()) and np.array_equal(blk(2, 3), -m["Xi_e"].toarray())
    assert np.array_equal(blk(3, 1), -m["Eps_f"].toarray()) and np.array_equal(blk(3, 2), -m["Xi_f"].toarray())
    for (i, j) in [(0, 0), (0, 3), (1, 1), (1, 2), (2, 1), (2, 2), (3, 0), (3, 3)]:
        assert not blk(i, j).any()


def test_nfgep_hermitian_for_real_phi():
    dims = (2, 3, 3)
    tab, _, _ = setup(dims, KS[0])
    svd = fc.svd_curl(tab)
    n = int(np.prod(dims))
    rng = np.random.default_rng(5)
    zeta = rng.standard_normal(24 * n) + 1j * rng.standard_normal(24 * n)
    Phi = rng.uniform(1, 3, 24 * n)
    A = dense(lambda v: fc.nfgep_A_matvec(v, tab, svd, np.conj(zeta), zeta, Phi), 32 * n)
    assert np.abs(A - A.conj().T).max() < 1e-12 * np.abs(A).max()
    # scalar coefficients act like constant diagonals
    x = rng.standard_normal(32 * n) + 1j * rng.standard_normal(32 * n)
    y1 = fc.nfgep_A_matvec(x, tab, svd, 0.3 - 0.2j, 0.3 + 0.2j, 2.0)
    y2 = fc.nfgep_A_matvec(x, tab, svd, np.full(24 * n, 0.3 - 0.2j), np.full(24 * n, 0.3 + 0.2j), np.full(24 * n, 2.0))
    assert np.allclose(y1, y2, rtol=1e-14, atol=0)

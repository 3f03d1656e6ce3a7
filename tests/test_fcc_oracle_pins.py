"""Pins of the FCC skewed-grid oracle (oracle/fcc_oracle.py) by what the mathematics
fixes, independent of the restated reshapes:

* T is unitary and TH is its adjoint;
* T equals the closed-form triple sum of twisted inverse DFTs (einsum with explicit
  DFT matrices, no reshapes);
* the explicit sparse difference matrices of Construct_GEP_mtx_A.m are diagonalised
  by T with the Lambda's of SVDSingleCurl_FCC.m (two reference files agreeing);
* each curl block factors as C_v (I_3 (x) T) Q_v = (I_3 (x) T) P_v diag(Sigma_v), with
  Q_v, P_v orthonormal;
* the 36 (tensor entry, index set) pairs of the parameter matrices follow from where the
  four staggered Yee grids sample each component, and equal the table read from the
  reference file;
* with the mtx_data that GEP B fixes (nfgep_data), compute_eigenvector is a solve with that B
  (the reference's own commented check), the NFGEP spectrum is the nonzero spectrum of the
  explicit (GEP A, GEP B), and mapped NFGEP eigenvectors solve the GEP.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import fcc_oracle as fc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

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


def _group_shifts():
    """Half-cell shift of each of the four Yee grids, read off the curl variants
    (SVDSingleCurl_FCC.m:33-36): a variant whose Lambda_d is -conj(Lambda_d) takes backward
    differences along d, i.e. its grid sits half a cell further along d."""
    L = (1 + 2j, 3 + 4j, 5 + 6j)
    tab = {k: np.array([v]) for k, v in zip(("Lambda_x", "Lambda_y", "Lambda_z"), L)}
    shifts = []
    for var in fc.variants(tab):
        sh = []
        for d in range(3):
            if var[d][0] == L[d]:
                sh.append(0.0)
            else:
                assert var[d][0] == -np.conj(L[d])
                sh.append(0.5)
        shifts.append(np.array(sh))
    return shifts


def _derived_param_table():
    """(row block, col block) -> (tensor row, tensor col, index set) from Yee-grid geometry alone:
    block 3g + a is component a of grid g; E_a sits on a-edges, H_a on a-faces, shifted by the
    grid's half-cell offset.  The material couples components sampled at one point, and that
    point's kind names the index set (edges X/Y/Z or cube centres for the E matrices, faces
    X/Y/Z or vertices for the H matrices; mtx_Eps_Mu_Zeta_Xi.m:23-35)."""
    sh = _group_shifts()
    e = 0.5 * np.eye(3)
    locE = lambda b: tuple(np.mod(e[b % 3] + sh[b // 3], 1.0))  # noqa: E731
    locH = lambda b: tuple(np.mod(0.5 - e[b % 3] + sh[b // 3], 1.0))  # noqa: E731
    edge = [(0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5), (0.5, 0.5, 0.5)]
    face = [(0, 0.5, 0.5), (0.5, 0, 0.5), (0.5, 0.5, 0), (0, 0, 0)]
    table = {}
    for rb in range(12):
        for cb in range(12):
            if locE(rb) == locE(cb):
                s = edge.index(locE(rb))
                assert face.index(locH(rb)) == s and locH(rb) == locH(cb)
                table[(rb, cb)] = (rb % 3, cb % 3, s)
    return table


def test_parameter_table_from_yee_geometry():
    """Every one of the 36 (tensor entry, index set) pairs of Construct_parameter_mtx.m:13-312
    follows from where the four staggered grids sample each field component."""
    table = _derived_param_table()
    assert len(table) == 36
    assert {(rb, cb): (a, b, s) for rb, cb, a, b, s in fc.PARAM_BLOCKS} == table


def test_parameter_table_matches_reference_text():
    """PARAM_BLOCKS equals, entry by entry and in order, the table read mechanically from the
    reference file by tests/golden/extract_param_blocks.py."""
    ref = json.load(open(os.path.join(GOLDEN, "param_blocks.json")))
    assert [tuple(e) for e in ref["entries"]] == list(fc.PARAM_BLOCKS)


def test_parameter_matrix_entries():
    """parameter_mtx places T_in[a, b] (inside the entry's index set) or T_air[a, b] at every
    (row block, col block, point) the geometry names, and nothing else."""
    rng = np.random.default_rng(9)
    n = 30
    sets = [rng.random(n) < 0.4 for _ in range(4)]
    Tin = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    Tair = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    M = fc.parameter_mtx(n, sets, Tin, Tair).toarray().reshape(12, n, 12, n)
    want = np.zeros_like(M)
    for (rb, cb), (a, b, s) in _derived_param_table().items():
        want[rb, np.arange(n), cb, np.arange(n)] = np.where(sets[s], Tin[a, b], Tair[a, b])
    assert np.array_equal(M, want)


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


# ---------------------------------------------------------------------------
# NFGEP against the explicit GEP (Construct_GEP_mtx_A.m, Construct_GEP_mtx_B.m)
# ---------------------------------------------------------------------------

def fcc_case(dims, k, seed=11, gamma=0.7, eps_in=(13.0, 13.0, 13.0)):
    """Random inside/air index sets; inside eps = diag(eps_in), mu = I, zeta = -i gamma I,
    xi = i gamma I (BASELINE's bi-isotropic inclusions), vacuum outside."""
    tab, (a1, a2, a3), delta = setup(dims, k)
    n = int(np.prod(dims))
    rng = np.random.default_rng(seed)
    es = [rng.random(n) < 0.5 for _ in range(4)]
    fs = [rng.random(n) < 0.5 for _ in range(4)]
    I3 = np.eye(3)
    tin = [np.diag(eps_in).astype(complex), I3, -1j * gamma * I3, 1j * gamma * I3]
    tair = [I3, I3, 0 * I3, 0 * I3]
    m = fc.eps_mu_zeta_xi(n, es, fs, tin, tair)
    blocks, _ = fc.curl_blocks(*dims, k, a1, a2, a3, delta)
    A, _ = fc.gep_A(blocks)
    return tab, fc.svd_curl(tab), m, A, fc.gep_B(m), n


def similarity(n):
    """R, P with R A P = A for GEP A (blocks scaled by 1 and -i*i) and R B P of the NFGEP form."""
    R = sp.block_diag([sp.identity(24 * n), -1j * sp.identity(24 * n)]).tocsr()
    P = sp.block_diag([sp.identity(24 * n), 1j * sp.identity(24 * n)]).tocsr()
    return R, P


@pytest.mark.parametrize("dims", [(2, 3, 3), (4, 3, 3)])
def test_eigenvector_map_is_gep_B_solve(dims):
    """Compute_eigenvector.m:27,41: ev equals GEP_mtx_B \\ [P_r u; Q_r v] for the B whose
    similarity transform the mtx_data comes from."""
    tab, svd, m, A, B, n = fcc_case(dims, KS[0])
    xi, zeta, Phi = fc.nfgep_data(m)
    R, P = similarity(n)
    assert abs(R @ A @ P - A).max() == 0
    rng = np.random.default_rng(2)
    x = rng.standard_normal(32 * n) + 1j * rng.standard_normal(32 * n)
    ev = fc.compute_eigenvector(x, tab, svd, xi, zeta, Phi, coupling=1)
    Tm = dense(lambda v: fc.T(v, tab), n)
    I24T = sp.kron(sp.identity(24), sp.csr_matrix(Tm))
    t = np.concatenate([I24T @ (svd["Lambda_Pr"] @ x[:16 * n]), I24T @ (svd["Lambda_Qr"] @ x[16 * n:])])
    ref = spl.spsolve((R @ B @ P).tocsc(), t)
    assert np.linalg.norm(ev - ref) <= 1e-12 * np.linalg.norm(ref)


def nfgep_dense(tab, svd, data, n, coupling=1):
    NF = dense(lambda v: fc.nfgep_A_matvec(v, tab, svd, *data, coupling=coupling), 32 * n)
    AF = dense(lambda v: fc.fncA(v, svd), 32 * n)
    return NF, AF


def match(w_small, w_big):
    """largest relative distance from each of w_small to the nearest of w_big"""
    return max(np.min(np.abs(w_big - w)) / abs(w) for w in w_small)


@pytest.mark.parametrize("ki,eps_in", [(0, (13.0, 13.0, 13.0)), (1, (12.0, 13.0, 14.0))])
def test_nfgep_spectrum_is_gep_spectrum(ki, eps_in):
    """The nonzero eigenvalues of the explicit (GEP A, GEP B), 48n of them less the 16n of A's
    null space, are exactly the NFGEP eigenvalues (mtx_prod_vec_A_NFGEP z = lambda fncA z)."""
    dims = (2, 3, 3)
    tab, svd, m, A, B, n = fcc_case(dims, KS[ki], eps_in=eps_in)
    NF, AF = nfgep_dense(tab, svd, fc.nfgep_data(m), n)
    w_nf = sl.eigvals(NF, AF)
    w_gep = sl.eigvals(A.toarray(), B.toarray())
    w_gep = w_gep[np.abs(w_gep) > 1e-9 * np.abs(w_gep).max()]
    assert w_nf.size == w_gep.size == 32 * n
    assert match(w_nf, w_gep) < 1e-10 and match(w_gep, w_nf) < 1e-10


def test_nfgep_eigenvectors_solve_the_gep():
    """For NFGEP eigenpairs (lambda, z), x = diag(I, iI) Compute_eigenvector(z) satisfies
    GEP_A x = lambda GEP_B x."""
    dims = (2, 3, 3)
    tab, svd, m, A, B, n = fcc_case(dims, KS[1], gamma=1.1)
    data = fc.nfgep_data(m)
    NF, AF = nfgep_dense(tab, svd, data, n)
    w, Z = sl.eig(NF, AF)
    _, P = similarity(n)
    order = np.argsort(np.abs(w))
    for i in list(order[:4]) + list(order[-2:]):
        x = P @ fc.compute_eigenvector(Z[:, i], tab, svd, *data, coupling=1)
        res = np.linalg.norm(A @ x - w[i] * (B @ x)) / (np.linalg.norm(x) * abs(w[i]))
        assert res < 1e-11, (i, w[i], res)


def test_nfgep_pin_detects_planted_mistakes():
    """A swap of Lambda_Pr / Lambda_Qr, of xi_d / zeta_d, or the diagonal coupling instead of the
    half coupling, planted in the oracle's inputs, moves the spectrum off the GEP's."""
    dims = (2, 3, 3)
    tab, svd, m, A, B, n = fcc_case(dims, KS[0])
    xi, zeta, Phi = fc.nfgep_data(m)
    w_gep = sl.eigvals(A.toarray(), B.toarray())
    w_gep = w_gep[np.abs(w_gep) > 1e-9 * np.abs(w_gep).max()]
    swapped = dict(svd, Lambda_Pr=svd["Lambda_Qr"], Lambda_Qr=svd["Lambda_Pr"])
    for s_, data, coupling in [(swapped, (xi, zeta, Phi), 1), (svd, (zeta, xi, Phi), 1), (svd, (xi, zeta, Phi), 0)]:
        NF, AF = nfgep_dense(tab, s_, data, n, coupling)
        assert match(sl.eigvals(NF, AF), w_gep) > 1e-3


def test_nfgep_data_rejects_what_it_cannot_represent():
    dims = (2, 3, 3)
    n = int(np.prod(dims))
    rng = np.random.default_rng(1)
    es = [rng.random(n) < 0.5 for _ in range(4)]
    I3 = np.eye(3)
    aniso = np.array([[5, 2, 1], [2, 5, 2], [1, 2, 4]], complex)
    with pytest.raises(ValueError):
        fc.nfgep_data(fc.eps_mu_zeta_xi(n, es, es, [aniso, I3, 0 * I3, 0 * I3], [I3, I3, 0 * I3, 0 * I3]))
    with pytest.raises(ValueError):
        fc.nfgep_data(fc.eps_mu_zeta_xi(n, es, es, [I3, 2 * I3, 0 * I3, 0 * I3], [I3, I3, 0 * I3, 0 * I3]))

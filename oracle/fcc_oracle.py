"""CPU oracle for the FCC skewed-grid operators (NumPy/SciPy FP64 restatement).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports, links or calls this
module; ``tests/`` use it as the checker of the device FCC operators.

Restates, from the text of the reference files (paths relative to the reference
root, MATLAB column-major reshapes written out with ``order="F"``):

* ``fft_tables`` / ``svd_curl``   light_locked_Huang/SVDSingleCurl_FCC.m:1-193
* ``T`` / ``TH``                  light_locked_Huang/p_mtx_T_prod_vec_FCC.m:1-37,
                                  light_locked_Huang/p_mtx_TH_prod_vec_FCC.m:1-28
* ``gamma_1/2/3``, ``curl_blocks``, ``gep_A``
                                  light_locked_Huang/Construct_GEP_mtx_A.m:1-505
* ``parameter_mtx``               light_locked_Huang/Construct_parameter_mtx.m:1-318
* ``eps_mu_zeta_xi``              light_locked_Huang/mtx_Eps_Mu_Zeta_Xi.m:1-38
* ``gep_B``                       light_locked_Huang/Construct_GEP_mtx_B.m:1-69
* ``nfgep_A_matvec``              light_locked_Huang/mtx_prod_vec_A_NFGEP.m:1-52 (diagonal Phi)
* ``compute_eigenvector``         light_locked_Huang/Compute_eigenvector.m:1-43 (diagonal Phi)
* ``nfgep_data``, ``fncA``        mtx_data and the NFGEP right-hand operator, derived from
                                  Construct_GEP_mtx_A/B.m (the reference never builds them)

Pinning (tests/test_fcc_oracle_pins.py): T is unitary; T equals the closed-form
triple sum of the twisted inverse DFTs; the explicit sparse difference matrices
are diagonalised by T (Gamma_d T = delta_d T diag(Lambda_d), two independent
reference files agreeing); every curl block factors as
C_v (I_3 (x) T) Q_v = (I_3 (x) T) P_v diag(Sigma_v); the parameter matrices are
four permuted copies of pointwise 3x3 tensors (in/air by index set).
``nfgep_A_matvec`` / ``compute_eigenvector``: the reference never builds its ``mtx_data``;
``nfgep_data`` derives it from GEP B, and then (tests/test_fcc_oracle_pins.py)
compute_eigenvector equals a sparse solve with that B (the reference's own commented check),
the NFGEP spectrum equals the nonzero spectrum of the explicit (GEP A, GEP B), and the mapped
NFGEP eigenvectors are GEP eigenvectors.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

TWO_PI_I = 2j * np.pi


# --------------------------------------------------------------------------
# SVDSingleCurl_FCC.m
# --------------------------------------------------------------------------

def fft_tables(n1, n2, n3, vec_k, a1, a2, a3, delta):
    """construct_fft_mtx_1219 (SVDSingleCurl_FCC.m:76-137) + the Fz/Fy layouts (:27-31).

    Frequency index order is z fastest: p = l + n3*(nn + n2*m) for the x-, y-, z-
    frequencies m, nn, l."""
    k = np.asarray(vec_k, float)
    a1, a2, a3 = (np.asarray(a, float) for a in (a1, a2, a3))
    D_kx = np.exp(TWO_PI_I * (k @ a1) / n1 * np.arange(n1))
    D_jx = np.empty((n2, n1), complex)
    for jj in range(n1):
        D_jx[:, jj] = np.exp(TWO_PI_I * (k @ (a2 - 0.5 * a1) - 0.5 * jj) / n2 * np.arange(n2))
    D_jell = np.empty((n3, n2, n1), complex)
    dp = a3 - (a1 + a2) / 3
    for jj in range(n1):
        for ll in range(n2):
            D_jell[:, ll, jj] = np.exp(TWO_PI_I * (k @ dp - (jj + ll) / 3) / n3 * np.arange(n3))
    lam_x = (np.exp(TWO_PI_I * (np.arange(n1) + k @ a1) / n1) - 1) / delta[0]
    Lambda_x = np.kron(lam_x, np.ones(n2 * n3))
    lam_y = np.empty(n1 * n2, complex)
    nk = np.linalg.norm(k)
    for m in range(n1):
        # at k = 0 the reference drops the -(m-1)/2 shift (SVDSingleCurl_FCC.m:111-114)
        dp2 = a2 - 0.5 * a1 if nk == 0 else a2 - 0.5 * (a1 + m / nk ** 2 * k)
        lam_y[m * n2:(m + 1) * n2] = (np.exp(TWO_PI_I * (np.arange(n2) + k @ dp2) / n2) - 1) / delta[1]
    Lambda_y = np.kron(lam_y, np.ones(n3))
    kd = k @ dp
    m, nn, l = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    Lambda_z = ((np.exp(TWO_PI_I * (l + kd - (m + nn) / 3) / n3) - 1) / delta[2]).ravel()
    Fz = np.empty((n3, n1 * n2), complex)
    for ii in range(n1):
        Fz[:, ii * n2:(ii + 1) * n2] = D_jell[:, :, ii]
    Fy = np.kron(np.ones((1, n3)), D_jx)
    return dict(n=(n1, n2, n3), D_kx=D_kx, D_jx=D_jx, D_jell=D_jell, Fz=Fz, Fy=Fy,
                Lambda_x=Lambda_x, Lambda_y=Lambda_y, Lambda_z=Lambda_z)


def lambda_dat(Lx, Ly, Lz):
    """construct_Lambda_dat (SVDSingleCurl_FCC.m:140-193): Lambda_q and the pointwise
    3-vectors of Q1, Q2 (returned as (3, n) arrays)."""
    Lq = (np.conj(Lx) * Lx + np.conj(Ly) * Ly + np.conj(Lz) * Lz).real
    Lp = Lx + Ly + Lz
    Lp = (Lp * np.conj(Lp)).real
    pq = 3 * Lq - Lp
    q2pq = np.sqrt(Lq * pq)
    pqs = np.sqrt(pq)
    Q2 = np.stack([(np.conj(Lz) - np.conj(Ly)) / pqs,
                   (np.conj(Lx) - np.conj(Lz)) / pqs,
                   (np.conj(Ly) - np.conj(Lx)) / pqs])
    Q1 = np.stack([np.conj(Lz) * (Lz - Lx) - np.conj(Ly) * (Lx - Ly),
                   np.conj(Lx) * (Lx - Ly) - np.conj(Lz) * (Ly - Lz),
                   np.conj(Ly) * (Ly - Lz) - np.conj(Lx) * (Lz - Lx)]) / q2pq
    return Lq, Q1, Q2


def variants(tab):
    """The four sign/conjugation variants of (Lambda_x, Lambda_y, Lambda_z) (SVDSingleCurl_FCC.m:33-36)."""
    Lx, Ly, Lz = tab["Lambda_x"], tab["Lambda_y"], tab["Lambda_z"]
    c = np.conj
    return [(Lx, Ly, Lz), (-c(Lx), -c(Ly), Lz), (-c(Lx), Ly, -c(Lz)), (Lx, -c(Ly), -c(Lz))]


def _pointwise(vecs3):
    """(3, n) pointwise column -> sparse 3n x n with vecs3[c][p] at (c*n + p, p)."""
    n = vecs3.shape[1]
    rows = (np.arange(3)[:, None] * n + np.arange(n)[None, :]).ravel()
    cols = np.tile(np.arange(n), 3)
    return sp.csr_matrix((vecs3.ravel(), (rows, cols)), shape=(3 * n, n))


def svd_curl(tab):
    """SVD_curl of SVDSingleCurl_FCC.m:38-72: Lambda_Qr, Lambda_Pr (24n x 16n), Sigma_r (8n),
    and the per-variant pointwise (Q1, Q2) pairs."""
    n = tab["Lambda_x"].size
    LQ, LP, sig, pairs = [], [], [], []
    for (Lx, Ly, Lz) in variants(tab):
        Lq, Q1, Q2 = lambda_dat(Lx, Ly, Lz)
        pairs.append((Q1, Q2))
        LQ.append(sp.hstack([_pointwise(Q1), _pointwise(Q2)]))
        LP.append(sp.hstack([_pointwise(-np.conj(Q2)), _pointwise(np.conj(Q1))]))
        sig += [np.sqrt(Lq), np.sqrt(Lq)]
    LQ = sp.block_diag(LQ)
    LP = sp.block_diag(LP)
    return dict(Lambda_Qr=sp.block_diag([LQ, LP]).tocsr(), Lambda_Pr=sp.block_diag([LP, LQ]).tocsr(),
                Sigma_r=np.concatenate(sig), pairs=pairs, n=n)


# --------------------------------------------------------------------------
# p_mtx_T_prod_vec_FCC.m / p_mtx_TH_prod_vec_FCC.m
# --------------------------------------------------------------------------

def T(x, tab):
    """T x (p_mtx_T_prod_vec_FCC.m:9-34): input z-fastest frequency order, output x-fastest grid order."""
    n1, n2, n3 = tab["n"]
    Q = x.reshape(n3, n1 * n2, order="F")
    Q = tab["Fz"] * np.fft.ifft(Q, axis=0)
    Qy = np.empty((n2, n1 * n3), complex)
    for ii in range(n1):
        Qy[:, ii * n3:(ii + 1) * n3] = Q[:, ii * n2:(ii + 1) * n2].T
    Qy = np.fft.ifft(Qy, axis=0)
    idx_col = np.concatenate([ii + n3 * np.arange(n1) for ii in range(n3)])
    Qy = tab["Fy"] * Qy[:, idx_col]
    Qx = np.empty((n1, n2 * n3), complex)
    for ii in range(n3):
        Qx[:, ii * n2:(ii + 1) * n2] = Qy[:, ii * n1:(ii + 1) * n1].T
    out = tab["D_kx"][:, None] * np.fft.ifft(Qx, axis=0)
    return np.sqrt(n1 * n2 * n3) * out.reshape(-1, order="F")


def TH(x, tab):
    """T^H x (p_mtx_TH_prod_vec_FCC.m:10-26) with the conjugated tables (SVDSingleCurl_FCC.m:17-23)."""
    n1, n2, n3 = tab["n"]
    Px = np.fft.fft(np.conj(tab["D_kx"])[:, None] * x.reshape(n1, n3 * n2, order="F"), axis=0)
    Py = np.empty((n2, n1 * n3), complex)
    for ii in range(n1):
        Py[:, ii * n3:(ii + 1) * n3] = np.conj(tab["D_jx"][:, ii])[:, None] * Px[ii, :].reshape(n2, n3, order="F")
    Py = np.fft.fft(Py, axis=0)
    Pz = np.empty((n3, n1 * n2), complex)
    for ii in range(n1):
        Pz[:, ii * n2:(ii + 1) * n2] = np.conj(tab["D_jell"][:, :, ii]) * Py[:, ii * n3:(ii + 1) * n3].T
    return np.fft.fft(Pz, axis=0).reshape(-1, order="F") / np.sqrt(n1 * n2 * n3)


# --------------------------------------------------------------------------
# Construct_GEP_mtx_A.m: the explicit skewed-grid difference matrices
# --------------------------------------------------------------------------

def _g2_entries(n1, n2, a1, a2, k):
    """Generate_mtx_G_2 (Construct_GEP_mtx_A.m:425-461): (rows, cols, vals) 0-based, in the
    reference's entry order (diagonal -1, +1 forward y neighbour, skewed wrap of the last y row)."""
    rows, cols, vals = [], [], []
    for jj in range(n2 - 1):
        ii = jj * n1
        rows += list(range(ii, ii + n1)); cols += list(range(ii, ii + n1)); vals += [-1.0] * n1
        rows += list(range(ii, ii + n1)); cols += list(range(ii + n1, ii + 2 * n1)); vals += [1.0] * n1
    ii = (n2 - 1) * n1
    rows += list(range(ii, ii + n1)); cols += list(range(ii, ii + n1)); vals += [-1.0] * n1
    t1 = TWO_PI_I * (k @ a1)
    t2 = TWO_PI_I * (k @ a2)
    h = n1 // 2
    rows += list(range(ii, ii + h)); cols += list(range(h, n1)); vals += [np.exp(t2 - t1)] * h
    rows += list(range(ii + h, ii + n1)); cols += list(range(0, h)); vals += [np.exp(t2)] * h
    return np.array(rows), np.array(cols), np.array(vals, complex)


def gamma_1(n1, n2, n3, a1, k):
    """Gamma_1 (Construct_GEP_mtx_A.m:394-418): x forward difference with the e^{2 pi i k.a1} wrap."""
    G = -np.eye(n1, dtype=complex)
    G[n1 - 1, 0] = np.exp(TWO_PI_I * (k @ a1))
    for ii in range(1, n1):
        G[ii - 1, ii] = 1
    return sp.kron(sp.identity(n2 * n3), sp.csr_matrix(G)).tocsr()


def gamma_2(n1, n2, n3, a1, a2, k):
    """Gamma_2 (Construct_GEP_mtx_A.m:420-423): Generate_mtx_G_2 repeated over z planes."""
    r, c, v = _g2_entries(n1, n2, a1, a2, k)
    G = sp.csr_matrix((v, (r, c)), shape=(n1 * n2, n1 * n2))
    return sp.kron(sp.identity(n3), G).tocsr()


def gamma_3(n1, n2, n3, a1, a2, a3, k):
    """Gamma_3 (Construct_GEP_mtx_A.m:463-505): z forward difference; the last plane wraps to the
    first with a y shift of n2/3 planes and, for y >= n2/3, the x half-shift of Gamma_2."""
    n12 = n1 * n2
    n = n12 * n3
    rows, cols, vals = [], [], []
    for jj in range(n3 - 1):
        ii = jj * n12
        rows += list(range(ii, ii + n12)); cols += list(range(ii, ii + n12)); vals += [-1.0] * n12
        rows += list(range(ii, ii + n12)); cols += list(range(ii + n12, ii + 2 * n12)); vals += [1.0] * n12
    ii = (n3 - 1) * n12
    rows += list(range(ii, ii + n12)); cols += list(range(ii, ii + n12)); vals += [-1.0] * n12
    t2 = TWO_PI_I * (k @ a2)
    t3 = TWO_PI_I * (k @ a3)
    third = n12 // 3
    rows += list(range(ii, ii + third)); cols += list(range(2 * third, n12)); vals += [np.exp(t3 - t2)] * third
    _, _, gv = _g2_entries(n1, n2, a1, a2, k)
    h = n1 // 2
    tail1 = np.exp(t3 - t2) * gv[2 * n12 - n1:2 * n12 - h]
    tail2 = np.exp(t3 - t2) * gv[2 * n12 - h:2 * n12]
    r0 = n - 2 * third
    c0 = h
    for _ in range(2 * n2 // 3):
        rows += list(range(r0, r0 + h)); cols += list(range(c0, c0 + h)); vals += list(tail1)
        r0 += h; c0 -= h
        rows += list(range(r0, r0 + h)); cols += list(range(c0, c0 + h)); vals += list(tail2)
        r0 += h; c0 += 3 * h
    return sp.csr_matrix((np.array(vals, complex), (rows, cols)), shape=(n, n))


def curl_blocks(n1, n2, n3, k, a1, a2, a3, delta):
    """The four 3n x 3n curl blocks of mtx_C (Construct_GEP_mtx_A.m:8-312), scaled by 1/delta."""
    k = np.asarray(k, float)
    G1 = gamma_1(n1, n2, n3, np.asarray(a1, float), k) / delta[0]
    G2 = gamma_2(n1, n2, n3, np.asarray(a1, float), np.asarray(a2, float), k) / delta[1]
    G3 = gamma_3(n1, n2, n3, np.asarray(a1, float), np.asarray(a2, float), np.asarray(a3, float), k) / delta[2]
    H = lambda M: M.conj().T  # noqa: E731

    def curl(g1, g2, g3):
        return sp.bmat([[None, -g3, g2], [g3, None, -g1], [-g2, g1, None]]).tocsr()

    # the variants' (G1, G2, G3) follow (Lx,Ly,Lz), (-Lx*,-Ly*,Lz), (-Lx*,Ly,-Lz*), (Lx,-Ly*,-Lz*)
    return [curl(G1, G2, G3), curl(-H(G1), -H(G2), G3), curl(-H(G1), G2, -H(G3)), curl(G1, -H(G2), -H(G3))], \
        (G1, G2, G3)


def gep_A(blocks):
    """GEP_mtx_A = blkdiag(C, C^H, C^H, C) with C = blkdiag(C_1..C_4) (Construct_GEP_mtx_A.m:314-320)."""
    C = sp.block_diag(blocks).tocsr()
    return sp.block_diag([C, C.conj().T, C.conj().T, C]).tocsr(), C


# --------------------------------------------------------------------------
# Construct_parameter_mtx.m / mtx_Eps_Mu_Zeta_Xi.m / Construct_GEP_mtx_B.m
# --------------------------------------------------------------------------

# (row block, col block, tensor row a, tensor col b, index set) for the 36 n-diagonals of
# Construct_parameter_mtx.m:13-312; set 0/1/2 = face_edge.X/Y/Z, 3 = vertex_Cubic.
PARAM_BLOCKS = [
    (0, 0, 0, 0, 0), (1, 1, 1, 1, 1), (2, 2, 2, 2, 2), (0, 4, 0, 1, 0), (1, 3, 1, 0, 1), (0, 8, 0, 2, 0),
    (2, 6, 2, 0, 2), (1, 11, 1, 2, 1), (2, 10, 2, 1, 2), (3, 1, 0, 1, 1), (4, 0, 1, 0, 0), (3, 3, 0, 0, 1),
    (4, 4, 1, 1, 0), (5, 5, 2, 2, 3), (4, 8, 1, 2, 0), (5, 7, 2, 1, 3), (3, 11, 0, 2, 1), (5, 9, 2, 0, 3),
    (6, 2, 0, 2, 2), (8, 0, 2, 0, 0), (7, 5, 1, 2, 3), (8, 4, 2, 1, 0), (6, 6, 0, 0, 2), (7, 7, 1, 1, 3),
    (8, 8, 2, 2, 0), (6, 10, 0, 1, 2), (7, 9, 1, 0, 3), (10, 2, 1, 2, 2), (11, 1, 2, 1, 1), (9, 5, 0, 2, 3),
    (11, 3, 2, 0, 1), (9, 7, 0, 1, 3), (10, 6, 1, 0, 2), (9, 9, 0, 0, 3), (10, 10, 1, 1, 2), (11, 11, 2, 2, 1),
]


def parameter_mtx(n, sets, T_in, T_air):
    """Construct_parameter_mtx.m: 12n x 12n sparse with 36 n-diagonals.  ``sets`` = four boolean
    masks (X, Y, Z, vertex/Cubic) of length n; T_in / T_air 3x3."""
    rows, cols, vals = [], [], []
    ar = np.arange(n)
    for (rb, cb, a, b, s) in PARAM_BLOCKS:
        v = np.where(sets[s], T_in[a, b], T_air[a, b]).astype(complex)
        rows.append(rb * n + ar); cols.append(cb * n + ar); vals.append(v)
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(12 * n, 12 * n))
    M.eliminate_zeros()  # the reference re-sparsifies after find (Construct_parameter_mtx.m:314-315)
    return M


def eps_mu_zeta_xi(n, edge_sets, face_sets, tensors_in, tensors_air):
    """mtx_Eps_Mu_Zeta_Xi.m:23-35: edge matrices use (edge X/Y/Z, Cubic), face matrices use
    (face X/Y/Z, vertex).  tensors_* = (eps, mu, zeta, xi) 3x3.  Returns the dict of the 8 matrices."""
    out = {}
    for name, Ti, Ta in zip(("Eps", "Mu", "Zeta", "Xi"), tensors_in, tensors_air):
        out[name + "_e"] = parameter_mtx(n, edge_sets, Ti, Ta)
        out[name + "_f"] = parameter_mtx(n, face_sets, Ti, Ta)
    return out


def gep_B(m):
    """Construct_GEP_mtx_B.m: [[0, Zf, Mf, 0], [Ze, 0, 0, Me], [-Ee, 0, 0, -Xe], [0, -Ef, -Xf, 0]] (48n)."""
    return sp.bmat([[None, m["Zeta_f"], m["Mu_f"], None],
                    [m["Zeta_e"], None, None, m["Mu_e"]],
                    [-m["Eps_e"], None, None, -m["Xi_e"]],
                    [None, -m["Eps_f"], -m["Xi_f"], None]]).tocsr()


# --------------------------------------------------------------------------
# mtx_prod_vec_A_NFGEP.m / Compute_eigenvector.m
# --------------------------------------------------------------------------

def _apply_chunks(f, v, n):
    return np.concatenate([f(v[i * n:(i + 1) * n]) for i in range(v.size // n)])


def _couple(c, v, coupling):
    """mtx_data.xi_d * v (or zeta_d * v) for a 24n x 24n matrix with one nonzero per row: the
    24n array ``c`` on the diagonal (coupling 0) or in the other 12n half (coupling 1)."""
    if coupling == 0:
        return c * v
    h = v.size // 2
    return c * np.concatenate([v[h:], v[:h]])


def nfgep_A_matvec(x, tab, svd, xi_d, zeta_d, Phi, coupling=0):
    """mtx_prod_vec_A_NFGEP.m:12-50 with mtx_data.Phi_is_diag = 'yes'.  xi_d, zeta_d: the nonzeros
    (24n, or scalars) of mtx_data.xi_d / zeta_d, diagonal (coupling 0) or coupling the 12n halves
    (coupling 1, the form ``nfgep_data`` derives from GEP B)."""
    n = svd["n"]
    tmp1 = _apply_chunks(lambda v: T(v, tab), svd["Lambda_Pr"] @ x[:16 * n], n)
    tmp2 = _apply_chunks(lambda v: T(v, tab), svd["Lambda_Qr"] @ x[16 * n:], n)
    vec1 = (_couple(xi_d, tmp1, coupling) + tmp2) / Phi
    tmp1 = tmp1 + _couple(zeta_d, vec1, coupling)
    out1 = svd["Lambda_Pr"].conj().T @ _apply_chunks(lambda v: TH(v, tab), tmp1, n)
    out2 = svd["Lambda_Qr"].conj().T @ _apply_chunks(lambda v: TH(v, tab), vec1, n)
    return np.concatenate([out1, out2])


def compute_eigenvector(x, tab, svd, xi_d, zeta_d, Phi, coupling=0):
    """Compute_eigenvector.m:9-37 (diagonal Phi): ev = -i [-vec1; tmp1 + zeta_d vec1] (48n)."""
    n = svd["n"]
    tmp1 = _apply_chunks(lambda v: T(v, tab), svd["Lambda_Pr"] @ x[:16 * n], n)
    tmp2 = _apply_chunks(lambda v: T(v, tab), svd["Lambda_Qr"] @ x[16 * n:], n)
    vec1 = (_couple(xi_d, tmp1, coupling) + tmp2) / Phi
    return -1j * np.concatenate([-vec1, tmp1 + _couple(zeta_d, vec1, coupling)])


def nfgep_data(m):
    """mtx_data (xi_d, zeta_d, Phi; coupling 1) from the eight parameter matrices of
    mtx_Eps_Mu_Zeta_Xi (mu = I, diagonal eps, zeta, xi).  With the 24n blocks of GEP B
    (Construct_GEP_mtx_B.m) B11 = [[0, Zf], [Ze, 0]], B12 = I, B21 = -diag(Ee, Ef),
    B22 = -[[0, Xe], [Xf, 0]], the similarity diag(I, -iI) B diag(I, iI) (GEP A unchanged) is
    i [[zeta_d, I], [-(Phi + xi_d zeta_d), -xi_d]], the B for which Compute_eigenvector returns
    B \\ [tmp1; tmp2] (its commented check, Compute_eigenvector.m:27,41):
    xi_d = i B22, zeta_d = -i B11, Phi = B21 - xi_d zeta_d."""
    diag = {}
    for k, M in m.items():
        M = sp.csr_matrix(M)
        if k.startswith("Mu"):
            if abs(M - sp.identity(M.shape[0])).max() != 0:
                raise ValueError("nfgep_data needs mu = I")
        elif abs(M - sp.diags(M.diagonal())).max() != 0:
            raise ValueError("nfgep_data needs diagonal eps, zeta, xi (Phi is not diagonal otherwise)")
        diag[k] = M.diagonal()
    Ee, Ef, Ze, Zf, Xe, Xf = (diag[k] for k in ("Eps_e", "Eps_f", "Zeta_e", "Zeta_f", "Xi_e", "Xi_f"))
    xi_d = np.concatenate([-1j * Xe, -1j * Xf])
    zeta_d = np.concatenate([-1j * Zf, -1j * Ze])
    Phi = np.concatenate([-Ee + Xe * Ze, -Ef + Xf * Zf])
    return xi_d, zeta_d, Phi


def fncA(x, svd):
    """The NFGEP's right-hand operator [i Sigma^-1 x_b; -i Sigma^-1 x_a] (32n), Sigma =
    [Sigma_r; Sigma_r] on each half: with GEP A = blkdiag(P_r Sigma Q_r^H, Q_r Sigma P_r^H)
    (Construct_GEP_mtx_A.m:314-320 through the SVD of SVDSingleCurl_FCC.m), the nonzero eigenvalues
    of (GEP A, GEP B) are those of mtx_prod_vec_A_NFGEP z = lambda fncA z (the FCC twin of
    fun_Lanczos/fncA.m:3-4)."""
    n = svd["n"]
    s = np.concatenate([svd["Sigma_r"], svd["Sigma_r"]])
    return np.concatenate([1j * x[16 * n:] / s, -1j * x[:16 * n] / s])


# lattice of the skewed FCC grid (Construct_idx_all_FCC.m:150-155): a1 = (1,0,0),
# a2 = (1/2, sqrt(3)/2, 0), a3 = (1/2, sqrt(3)/6, sqrt(6)/3), times the edge length
def fcc_lattice(edge_len=1.0):
    return (edge_len * np.array([1.0, 0.0, 0.0]), edge_len * np.array([0.5, np.sqrt(3) / 2, 0.0]),
            edge_len * np.array([0.5, np.sqrt(3) / 6, np.sqrt(6) / 3]))

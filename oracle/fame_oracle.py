"""CPU oracle: FP64 NumPy/SciPy restatement of the reference's curl-SVD + material matvec path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library and its
Python mirror) imports, links or calls this module.  It is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, where it is the checker and the CPU baseline, never the thing
measured.

The reference is MATLAB (FAME-GPU/3D-Bi-Anisotropic-PhC) and cannot run in this
pipeline, so every function below restates one ``.m`` file from its text, with
the same matrices built the same way (sparse Kronecker products, the same
sign / conjugation / normalisation conventions).  Citations are
``path:line`` relative to the reference root.

Index convention (MATLAB column-major ``reshape``): grid point
``p = i1 + n1*(i2 + n2*i3)``, i.e. x fastest.  A 3-component field is
``[c1 (N); c2 (N); c3 (N)]``; an NFSEP vector is ``[x_a (2N); x_b (2N)]``.

Pinning status (what checks each function; see ``tests/test_oracle_pins.py``):

* ``lambda_quasiperiodic`` / ``S_sDFT`` / ``DFT_mtx``: K_QPBC = V diag(lambda) V^H
  with V = diag(S) F (dense eig identity, SVD_Matrix_1D.m:49-59).
* ``svd_matrix_generate`` + ``MtxVecMul_FFT`` + ``Pr/Qr/Prs/Qrs``: explicit sparse
  curl C = Pr diag(Sigma_r) Qr^H and C^H = Qr diag(Sigma_r) Pr^H
  (Funhand_Generate_Lanzcos.m:218-219), orthonormality (…:213-214), dense DFT
  matrix product at tiny n.
* material interpolation (``interpolate_material``): nnz/row 9/25/20, vacuum
  gives N_int = M_int = I and J_int = 0, Hermiticity, and the vertex-factorised
  stencil ``material_apply_factored`` equals the sparse products.
* ``fncA/fncB``: vacuum fncB = I; Hermiticity; the NFSEP spectrum equals the
  non-zero spectrum of the original curl problem built only from the explicit
  curl (EvaluteError_Chiral_no_inv.m:43-44) by dense eig at tiny n.
* ``fncinvB``: ||B y - x|| <= tol ||x||.
* ``minres``: the k-th iterate equals the least-squares minimiser of ||b - A x||
  over the explicit Krylov space K_k(A, b) (brute force at tiny size), and the
  converged solution equals a dense solve.
* ``lanczos``: vacuum closed form omega = sqrt(Lambda_q) and dense ``eigh`` of the
  full GEP at tiny n.
* ``eigen_restoration`` / ``evaluate_error``: original-problem residual small for
  converged pairs.
"""
from __future__ import annotations

import os

import numpy as np
import scipy.fft
import scipy.sparse as sp

# host threads for the FFTs: the CPU share given by OMP_NUM_THREADS, at most the machine's cores
FFT_WORKERS = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16"))))

# --------------------------------------------------------------------------
# 1D quasi-periodic SVD pieces (subroutines/@DistOp)
# --------------------------------------------------------------------------


def theta_m(m, k):
    """DistOp.theta: (k-1)*pi/m   (subroutines/@DistOp/theta.m:2)."""
    return (np.asarray(k, dtype=float) - 1.0) * np.pi / m


def lambda_quasiperiodic(m, theta, delta):
    """Singular values of K_QPBC (subroutines/@DistOp/lambda_QuasiPeriodic.m:3).

    lambda(j) = (exp(i(2*theta(m,j) + theta/m)) - 1)/delta, j = 1..m.
    """
    j = np.arange(1, m + 1)
    return (np.exp(1j * (2.0 * theta_m(m, j) + theta / m)) - 1.0) / delta


def S_sDFT(m, theta):
    """Bloch twist exp(i*(0:m-1)*theta/m)   (subroutines/@DistOp/S_sDFT.m:2)."""
    return np.exp(1j * np.arange(m) * theta / m)


def K_QPBC(m, theta, delta):
    """Forward difference with quasi-periodic wrap (subroutines/@DistOp/K_QPBC.m:3)."""
    rows = np.r_[np.arange(m), np.arange(m)]
    cols = np.r_[np.arange(m), np.arange(1, m), 0]
    vals = np.r_[-np.ones(m), np.ones(m - 1), np.exp(1j * theta)].astype(complex)
    return sp.csr_matrix((vals, (rows, cols)), shape=(m, m)) / delta


def DFT_mtx(m):
    """Unitary DFT matrix sqrt(1/m) exp(2i*j*theta(m,k))  (subroutines/@DistOp/DFT_mtx.m:3)."""
    j = np.arange(m)[:, None]
    return np.sqrt(1.0 / m) * np.exp(2j * j * theta_m(m, np.arange(1, m + 1))[None, :])


def theta_from_k(wave_vec, lattice_vec):
    """theta_d = k . a_d   (band_structure_woodpile_Lanczos_n60_gamma0p2.m:49-56).

    ``lattice_vec`` holds the lattice vectors as COLUMNS (MATLAB
    ``opt.phys.lattice.lattice_vector``).  A zero k is nudged by +1e-4 in every
    component, as at :50-52.
    """
    k = np.asarray(wave_vec, dtype=float).copy()
    if np.linalg.norm(k) == 0:
        k = k + 1e-4
    A = np.asarray(lattice_vec, dtype=float)
    return np.array([k @ A[:, 0], k @ A[:, 1], k @ A[:, 2]])


# --------------------------------------------------------------------------
# 3D tables: SVD_Matrix_3D + SVD_Matrix_curl (QP on all axes: m_d = n_d)
# --------------------------------------------------------------------------


def svd_matrix_generate(dims, theta, delta):
    """Per-k tables of the null-space-free curl SVD.

    Restates SVD_Matrix_Generate.m:24-30 -> SVD_Matrix_1D.m:29-48 (QP branch),
    SVD_Matrix_3D.m:8-17,72-86,172-179 and SVD_Matrix_curl.m:29-32,45-67,74-75,88-89
    for quasi-periodic boundary conditions on all three axes (then m_d = n_d,
    Ntil = Nhat = 0, I_Psi = I_Phi = I and every factor is sDFT, so
    Psi == Phi).

    Returns a dict with the 1D singular values, the 3D diagonals Lx/Ly/Lz, the
    twist S (identical for all components and for Psi/Phi), the sparse Pi.Qr /
    Pi.Pr (3N x 2N) and Sigma_r (2N).
    """
    n1, n2, n3 = (int(d) for d in dims)
    N = n1 * n2 * n3
    lx = lambda_quasiperiodic(n1, theta[0], delta[0])
    ly = lambda_quasiperiodic(n2, theta[1], delta[1])
    lz = lambda_quasiperiodic(n3, theta[2], delta[2])
    # SVD_Matrix_3D.m:15-17   Lambda_3D.x = kron(I_m3, kron(I_m2, Lambda_1D.x)) ...
    Lx = np.kron(np.ones(n3), np.kron(np.ones(n2), lx))
    Ly = np.kron(np.ones(n3), np.kron(ly, np.ones(n1)))
    Lz = np.kron(lz, np.ones(n2 * n1))
    # SVD_Matrix_3D.m:172-179   S.Psi_c = kron(V.z.S, kron(V.y.S, U.x.S)) ...
    Sx, Sy, Sz = S_sDFT(n1, theta[0]), S_sDFT(n2, theta[1]), S_sDFT(n3, theta[2])
    S1 = np.kron(Sz, np.kron(Sy, Sx))
    S = np.concatenate([S1, S1, S1])

    dg = lambda v: sp.diags(v, 0, shape=(N, N), format="csr")  # noqa: E731
    O = sp.csr_matrix((N, N), dtype=complex)
    # SVD_Matrix_curl.m:29-32
    Lambda_r = sp.bmat([[O, -dg(Lz), dg(Ly)], [dg(Lz), O, -dg(Lx)], [-dg(Ly), dg(Lx), O]], format="csr")
    # SVD_Matrix_curl.m:45-53
    w = np.arange(1, 4, dtype=float)
    w = w / np.linalg.norm(w)
    Lambda_q = np.abs(Lx) ** 2 + np.abs(Ly) ** 2 + np.abs(Lz) ** 2
    Lambda_w = w[0] * Lx + w[1] * Ly + w[2] * Lz
    norm_Pi_0 = np.sqrt(Lambda_q)
    norm_Pi_1 = np.sqrt(Lambda_q - np.abs(Lambda_w) ** 2)
    # SVD_Matrix_curl.m:55-57
    Pi_0 = sp.vstack([dg(Lx), dg(Ly), dg(Lz)]) @ dg(1.0 / norm_Pi_0)
    kw = sp.vstack([w[0] * sp.identity(N), w[1] * sp.identity(N), w[2] * sp.identity(N)])
    Pi_1 = (-(Lambda_r.conj().T) @ kw @ dg(1.0 / norm_Pi_1)).tocsr()
    Pi_2 = ((Lambda_r.conj().T) @ Pi_1.conj() @ dg(1.0 / norm_Pi_0)).tocsr()
    # SVD_Matrix_curl.m:63-64, 88-89 (I_Psi = I_Phi = I, Ntil = 0)
    Qr = sp.hstack([Pi_1, Pi_2], format="csr")
    Pr = sp.hstack([-Pi_2.conj(), Pi_1.conj()], format="csr")
    # SVD_Matrix_curl.m:50,67,74
    Sigma_r = np.concatenate([norm_Pi_0, norm_Pi_0])
    D_Pi_1 = np.concatenate([Pi_1[c * N:(c + 1) * N, :].diagonal() for c in range(3)])
    D_Pi_2 = np.concatenate([Pi_2[c * N:(c + 1) * N, :].diagonal() for c in range(3)])
    return dict(dims=(n1, n2, n3), N=N, theta=np.asarray(theta, float), delta=np.asarray(delta, float),
                lx=lx, ly=ly, lz=lz, Lx=Lx, Ly=Ly, Lz=Lz, S=S, Lambda_r=Lambda_r,
                Lambda_q=Lambda_q, Pi_0=Pi_0, Pi_1=Pi_1, Pi_2=Pi_2, D_Pi_1=D_Pi_1, D_Pi_2=D_Pi_2,
                Qr=Qr, Pr=Pr, Sigma_r=Sigma_r)


# --------------------------------------------------------------------------
# Singular-vector matvecs (gMtxVecMul_FFT / MtxVecMul_FFT, QP -> pure DFT)
# --------------------------------------------------------------------------


def _cube(v, dims):
    n1, n2, n3 = dims
    return v.reshape(n3, n2, n1)  # MATLAB reshape(v,[n1 n2 n3]) with axes reversed


def MtxVecMul_FFT(vec_in, dims, istrans, S, D=None):
    """Psi / Phi ('notranspose') and Psi^H / Phi^H ('transpose').

    Restates subroutines/gMtxVecMul_FFT.m:6-61 for sDFT factors (P, R, P_idx are
    identities, D = 1):  'notranspose': y_c = S .* (ifftn(D.*x_c) * sqrt(N));
    'transpose': y_c = conj(D) .* fftn(conj(S).*x_c) / sqrt(N).
    """
    dims = tuple(int(d) for d in dims)
    N = dims[0] * dims[1] * dims[2]
    norm_c = np.sqrt(N)
    D = np.ones(3 * N) if D is None else D
    if istrans == "notranspose":
        x = D * vec_in
        out = [scipy.fft.ifftn(_cube(x[c * N:(c + 1) * N], dims), workers=FFT_WORKERS).ravel() * norm_c for c in range(3)]
        return S * np.concatenate(out)
    if istrans == "transpose":
        x = np.conj(S) * vec_in
        out = [scipy.fft.fftn(_cube(x[c * N:(c + 1) * N], dims), workers=FFT_WORKERS).ravel() / norm_c for c in range(3)]
        return np.conj(D) * np.concatenate(out)
    raise ValueError(istrans)


class Fnc:
    """Function-handle struct ``fnc`` (fun_Lanczos/Funhand_Generate_Lanzcos.m:38-41,79-82)."""

    def __init__(self, mtx_A):
        self.A = mtx_A
        dims, S = mtx_A["dims"], mtx_A["S"]
        self.Psis = lambda x: MtxVecMul_FFT(x, dims, "transpose", S)
        self.Phis = lambda x: MtxVecMul_FFT(x, dims, "transpose", S)
        self.Psi = lambda x: MtxVecMul_FFT(x, dims, "notranspose", S)
        self.Phi = lambda x: MtxVecMul_FFT(x, dims, "notranspose", S)
        Pr, Qr = mtx_A["Pr"], mtx_A["Qr"]
        PrH, QrH = Pr.conj().T.tocsr(), Qr.conj().T.tocsr()
        self.Pr = lambda x: self.Psi(Pr @ x)
        self.Qr = lambda x: self.Phi(Qr @ x)
        self.Prs = lambda y: PrH @ self.Psis(y)
        self.Qrs = lambda y: QrH @ self.Phis(y)


# --------------------------------------------------------------------------
# Explicit curl (DistOp.C), used only as an independent pin
# --------------------------------------------------------------------------


def curl_C(dims, delta, theta):
    """Explicit sparse curl for e (subroutines/@DistOp/C.m:3-5 with C_12..C_32.m, QP)."""
    n1, n2, n3 = dims
    I1, I2, I3 = sp.identity(n1), sp.identity(n2), sp.identity(n3)
    Kx = K_QPBC(n1, theta[0], delta[0])
    Ky = K_QPBC(n2, theta[1], delta[1])
    Kz = K_QPBC(n3, theta[2], delta[2])
    C12 = sp.kron(I3, sp.kron(Ky, I1))
    C13 = sp.kron(Kz, sp.kron(I2, I1))
    C21 = sp.kron(I3, sp.kron(I2, Kx))
    C23 = sp.kron(Kz, sp.kron(I2, I1))
    C31 = sp.kron(I3, sp.kron(I2, Kx))
    C32 = sp.kron(I3, sp.kron(Ky, I1))
    N = n1 * n2 * n3
    O = sp.csr_matrix((N, N), dtype=complex)
    return sp.bmat([[O, -C13, C12], [C23, O, -C21], [-C32, C31, O]], format="csr")


# --------------------------------------------------------------------------
# Material: inversion + covariant form, vertex tensors, interpolation
# --------------------------------------------------------------------------


def invert_constitutive(eps, xi, zeta, mu):
    """Inverse 6x6 constitutive blocks (InputData/BS_Gyroid/Physical_settings_woodpile.m:91-96).

    Returns (eps', xi', zeta', mu') = blocks of inv([eps xi; zeta mu]).
    """
    big = np.block([[np.asarray(eps, complex), np.asarray(xi, complex)],
                    [np.asarray(zeta, complex), np.asarray(mu, complex)]])
    b = np.linalg.inv(big)
    return b[:3, :3], b[:3, 3:], b[3:, :3], b[3:, 3:]


def covariant(X, A):
    """Covariant form A'*(X*A)   (Physical_settings_woodpile.m:184-194)."""
    A = np.asarray(A, float)
    return A.T @ (np.asarray(X, complex) @ A)


def material_tensors(eps, mu, xi, zeta, lattice_vec):
    """(N-slot, M-slot, J-slot) 3x3 tensors for one medium: inverse blocks, covariant.

    Slot map (Public/Material_Matrix_Generate_Mult_Point.m:28-31):
    N <- permittivity_cov (inverse (1:3,1:3)), M <- permeability_cov
    (inverse (4:6,4:6)), J <- magnetoelectric_Xi_cov (inverse (1:3,4:6)).
    """
    e_i, x_i, _z_i, m_i = invert_constitutive(eps, xi, zeta, mu)
    return covariant(e_i, lattice_vec), covariant(m_i, lattice_vec), covariant(x_i, lattice_vec)


def dist_material_matrix(N, par_out, par_in, inner_idx):
    """Per-vertex diagonal blocks A{i,j}   (Material_Matrix_Generate_Mult_Point.m:38-48).

    Returns an (N,3,3) array of vertex values (the diagonals of A{i,j}).
    """
    I_t = np.zeros(N)
    I_t[np.asarray(inner_idx, dtype=np.int64)] = 1.0
    par_out = np.asarray(par_out, complex)
    par_in = np.asarray(par_in, complex)
    return (par_in - par_out)[None, :, :] * I_t[:, None, None] + par_out[None, :, :]


def _shift_ops(dims, theta):
    """C_b1..3 with delta = 1 (subroutines/@DistOp/C_b1.m:2, C_b2.m, C_b3.m) and P = I_R1 = I."""
    n1, n2, n3 = dims
    I1, I2, I3 = sp.identity(n1), sp.identity(n2), sp.identity(n3)
    Cb = [sp.kron(I3, sp.kron(I2, K_QPBC(n1, theta[0], 1.0))).tocsr(),
          sp.kron(I3, sp.kron(K_QPBC(n2, theta[1], 1.0), I1)).tocsr(),
          sp.kron(K_QPBC(n3, theta[2], 1.0), sp.kron(I2, I1)).tocsr()]
    P = sp.identity(n1 * n2 * n3, dtype=complex, format="csr")
    return P, Cb


def _setdiff(i):
    return [d for d in range(3) if d != i]


def IntCovarFormN_vertex(A, P, C):
    """fun_no_inv/IntCovarFormN_vertex.m:3-14 (A: list of lists of sparse diagonals)."""
    tmp = [[None] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            if i != j:
                tmp[i][j] = 0.25 * (2 * P + C[i]) @ A[i][j] @ (2 * P + C[j].conj().T)
            else:
                tmp[i][j] = 0.5 * (A[i][j] + (P + C[i]) @ A[i][j] @ (P + C[j]).conj().T)
    return sp.bmat(tmp, format="csr")


def IntCovarFormM_vertex_new(A, I, C):
    """fun_no_inv/IntCovarFormM_vertex_new.m:3-19."""
    tmp = [[None] * 3 for _ in range(3)]
    for i in range(3):
        k = _setdiff(i)
        for j in range(3):
            if i == j:
                a = A[i][i]
                tmp[i][i] = 0.25 * (a + (I + C[k[0]]) @ a @ (I + C[k[0]].conj().T)
                                    + (I + C[k[1]]) @ a @ (I + C[k[1]].conj().T)
                                    + (I + C[k[0]]) @ (I + C[k[1]]) @ a @ (I + C[k[1]].conj().T) @ (I + C[k[0]].conj().T))
            else:
                l = _setdiff(j)
                tmp[i][j] = (1.0 / 16) * (2 * I + C[k[0]]) @ (2 * I + C[k[1]]) @ A[i][j] \
                    @ (2 * I + C[l[1]].conj().T) @ (2 * I + C[l[0]].conj().T)
    return sp.bmat(tmp, format="csr")


def IntCovarFormJ_vertex_new(A, I, C):
    """fun_no_inv/IntCovarFormJ_vertex_new.m:3-11."""
    tmp = [[None] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            k = _setdiff(j)
            tmp[i][j] = 0.125 * (2 * I + C[i]) @ A[i][j] @ (2 * I + C[k[1]].conj().T) @ (2 * I + C[k[0]].conj().T)
    return sp.bmat(tmp, format="csr")


def interpolate_material(dims, theta, N_vert, M_vert, J_vert):
    """Interpolated material matrices N_int, M_int, J_int.

    Restates fun_no_inv/Interpolation_Material_Matrix_Mult_Point_vertex.m:14-27:
    N_int and M_int are symmetrised as 1/2 (X + X'); J_int is not.
    ``*_vert`` are (N,3,3) per-vertex tensors (the diagonals of mtx_B.N/M/J).
    """
    N = int(np.prod(dims))
    P, Cb = _shift_ops(dims, theta)
    blk = lambda T: [[sp.diags(T[:, i, j], 0, shape=(N, N), format="csr") for j in range(3)] for i in range(3)]  # noqa: E731
    N_int = IntCovarFormN_vertex(blk(N_vert), P, Cb)
    N_int = (0.5 * (N_int + N_int.conj().T)).tocsr()
    M_int = IntCovarFormM_vertex_new(blk(M_vert), P, Cb)
    M_int = (0.5 * (M_int + M_int.conj().T)).tocsr()
    J_int = IntCovarFormJ_vertex_new(blk(J_vert), P, Cb).tocsr()
    return dict(N_int=N_int, M_int=M_int, J_int=J_int)


def material_apply_factored(dims, theta, N_vert, M_vert, J_vert, h, e):
    """The same operator as ``[M h + J^H e; J h + N e]`` written vertex-wise.

    Interpolation = gather to vertices, a pointwise 6x6 product with the
    (symmetrised) vertex tensors, scatter back, plus the lumped diagonal
    blocks.  With sigma_d = I + C_b,d (shift by +1 along d, Bloch wrap):
    E_d = (I + sigma_d)/2 (vertex -> edge d), F_d = (I + sigma_k1)(I + sigma_k2)/4
    (vertex -> face d, k = axes != d).  Then
      J = E_i xi_ij F_j^H,  N_ij = E_i epsS_ij E_j^H (i != j),
      M_ij = F_i muS_ij F_j^H (i != j),
      N_ii = 1/2 (epsS_ii + sigma_i epsS_ii sigma_i^H),
      M_ii = 1/4 sum over the 4 face vertices of muS_ii,
    where XS = (X + X^H)/2 per vertex (the symmetrisation of
    Interpolation_Material_Matrix_Mult_Point_vertex.m:22,24).  This is the
    form the device kernels evaluate; tests pin it to the sparse matrices.
    """
    n1, n2, n3 = dims
    N = n1 * n2 * n3
    c = lambda v: v.reshape(n3, n2, n1)  # noqa: E731
    axis = {0: 2, 1: 1, 2: 0}
    n_ax = {0: n1, 1: n2, 2: n3}

    def sig(u, d):  # (sigma_d u)(p) = u(p + e_d), wrap phase exp(i theta_d)
        out = np.roll(u, -1, axis=axis[d])
        sl = [slice(None)] * 3
        sl[axis[d]] = n_ax[d] - 1
        out[tuple(sl)] *= np.exp(1j * theta[d])
        return out

    def sigH(u, d):  # adjoint: u(p - e_d), wrap phase exp(-i theta_d)
        out = np.roll(u, 1, axis=axis[d])
        sl = [slice(None)] * 3
        sl[axis[d]] = 0
        out[tuple(sl)] *= np.exp(-1j * theta[d])
        return out

    E = lambda u, d: 0.5 * (u + sig(u, d))  # noqa: E731
    EH = lambda u, d: 0.5 * (u + sigH(u, d))  # noqa: E731

    def F(u, d):
        k0, k1 = _setdiff(d)
        return E(E(u, k0), k1)

    def FH(u, d):
        k0, k1 = _setdiff(d)
        return EH(EH(u, k0), k1)

    hf = [c(h[i * N:(i + 1) * N]) for i in range(3)]
    ef = [c(e[i * N:(i + 1) * N]) for i in range(3)]
    T = lambda A, i, j: c(A[:, i, j])  # noqa: E731
    epsS = lambda i, j: 0.5 * (T(N_vert, i, j) + np.conj(T(N_vert, j, i)))  # noqa: E731
    muS = lambda i, j: 0.5 * (T(M_vert, i, j) + np.conj(T(M_vert, j, i)))  # noqa: E731
    ve = [EH(ef[j], j) for j in range(3)]
    vh = [FH(hf[j], j) for j in range(3)]
    ye, yh = [], []
    for i in range(3):
        we = sum(epsS(i, j) * ve[j] for j in range(3) if j != i) + sum(T(J_vert, i, j) * vh[j] for j in range(3))
        wh = sum(muS(i, j) * vh[j] for j in range(3) if j != i) + sum(np.conj(T(J_vert, j, i)) * ve[j] for j in range(3))
        a = epsS(i, i)
        de = 0.5 * (a * ef[i] + sig(a * sigH(ef[i], i), i))
        k0, k1 = _setdiff(i)
        b = muS(i, i)
        sk0 = lambda u: sig(b * sigH(u, k0), k0)  # noqa: E731
        dh = 0.25 * (b * hf[i] + sk0(hf[i]) + sig(b * sigH(hf[i], k1), k1)
                     + sig(sig(b * sigH(sigH(hf[i], k1), k0), k0), k1))
        ye.append((E(we, i) + de).ravel())
        yh.append((F(wh, i) + dh).ravel())
    return np.concatenate(yh), np.concatenate(ye)


# --------------------------------------------------------------------------
# NFSEP operators (fun_Lanczos/fncA.m, fncB.m, fncinvB.m)
# --------------------------------------------------------------------------


def fncA(x, Sigma_r):
    """y = [i*inv_r.*x_b; -i*inv_r.*x_a]   (fun_Lanczos/fncA.m:3-4)."""
    h = x.shape[0] // 2
    inv_r = 1.0 / Sigma_r
    return np.concatenate([1j * inv_r * x[h:], -1j * inv_r * x[:h]])


def fncB(x, fnc, mtx_B):
    """y = [Prs(M h + J' e); Qrs(J h + N e)], h = Pr(x_a), e = Qr(x_b)  (fun_Lanczos/fncB.m:3-6)."""
    h2 = x.shape[0] // 2
    hh = fnc.Pr(x[:h2])
    ee = fnc.Qr(x[h2:])
    M, J, Nn = mtx_B["M_int"], mtx_B["J_int"], mtx_B["N_int"]
    y_h = M @ hh + J.conj().T @ ee
    y_e = J @ hh + Nn @ ee
    return np.concatenate([fnc.Prs(y_h), fnc.Qrs(y_e)])


def pcg(op, b, tol, maxit, x0=None):
    """Unpreconditioned conjugate gradients (MATLAB ``pcg`` as called by fun_Lanczos/fncinvB.m:4).

    Stops when ||b - A x|| <= tol*||b|| (recursive residual, confirmed by one
    explicit residual as MATLAB does).  Returns (x, flag, relres, iters).
    Iteration counts are not pinned to MATLAB's (its stagnation logic is
    closed); the solution is pinned by the residual bound.
    """
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - op(x) if x0 is not None else b.copy()
    nb = np.linalg.norm(b)
    if nb == 0:
        return np.zeros_like(b), 0, 0.0, 0
    p = r.copy()
    rho = np.vdot(r, r).real
    for it in range(1, maxit + 1):
        q = op(p)
        alpha = rho / np.vdot(p, q).real
        x = x + alpha * p
        r = r - alpha * q
        rn = np.linalg.norm(r)
        if rn <= tol * nb:
            rt = np.linalg.norm(b - op(x))
            if rt <= tol * nb:
                return x, 0, rt / nb, it
        rho_new = np.vdot(r, r).real
        p = r + (rho_new / rho) * p
        rho = rho_new
    return x, 1, np.linalg.norm(b - op(x)) / nb, maxit


def minres(op, b, tol, maxit):
    """Unpreconditioned MINRES (Paige & Saunders 1975) for a Hermitian operator.

    The SIRA inner solver (light_locked_Huang/SIRA_solve_residual_LS_GEP.m:50-53)
    calls ``minres1``, which the reference does not ship; this is the published
    algorithm: Lanczos tridiagonalisation, Givens QR of the tridiagonal, and the
    three-term update of the search directions w.  Stops when the recursive
    residual |eta_k| <= tol*||b||.  Returns (x, flag, relres, iters)."""
    n = b.shape[0]
    x = np.zeros(n, complex)
    beta1 = np.linalg.norm(b)
    if beta1 == 0:
        return x, 0, 0.0, 0
    v_old = np.zeros(n, complex)
    v = b / beta1
    w1 = np.zeros(n, complex)   # w_{k-1}
    w2 = np.zeros(n, complex)   # w_{k-2}
    beta = 0.0
    c1, s1, c2, s2 = 1.0, 0.0, 1.0, 0.0  # rotations k-1 and k-2
    eta = beta1
    for k in range(1, maxit + 1):
        p = op(v)
        alpha = np.vdot(v, p).real
        p = p - alpha * v - beta * v_old
        beta_new = np.linalg.norm(p)
        delta = c1 * alpha - c2 * s1 * beta      # rotated diagonal of column k
        rho2 = s1 * alpha + c2 * c1 * beta       # entry k-1
        rho3 = s2 * beta                         # entry k-2
        rho1 = np.hypot(delta, beta_new)
        c, s = (delta / rho1, beta_new / rho1) if rho1 > 0 else (1.0, 0.0)
        w = (v - rho3 * w2 - rho2 * w1) / rho1 if rho1 > 0 else np.zeros(n, complex)
        x = x + c * eta * w
        eta = -s * eta
        w2, w1 = w1, w
        c2, s2, c1, s1 = c1, s1, c, s
        v_old = v
        v = p / beta_new if beta_new > 1e-300 else np.zeros(n, complex)
        beta = beta_new
        if abs(eta) <= tol * beta1 or beta_new <= 1e-300 or rho1 == 0:
            return x, 0, abs(eta) / beta1, k
    return x, 1, abs(eta) / beta1, maxit


def fncinvB(x, B_op, tol, maxit):
    """y = B \\ x by pcg (fun_Lanczos/fncinvB.m:4)."""
    y, *_ = pcg(B_op, x, tol, maxit)
    return y


# --------------------------------------------------------------------------
# Lanczos (fun_Lanczos/Lanczos.m)
# --------------------------------------------------------------------------


def lanczos(fA, fB, fBinv, n, k, tol, maxiter, p, v0):
    """Thick-restart B-orthogonal Lanczos on B^{-1}A  (fun_Lanczos/Lanczos.m:1-156).

    Returns (ev, ew, iters) with ew the k largest eigenvalues of B^{-1}A by
    real part (descending), as the reference's ``[ev, ew, iter]``.
    """
    v = v0 / np.sqrt(abs(np.vdot(v0, fB(v0))))  # Lanczos.m:8
    eps = 2.2204e-16
    Q = np.zeros((n, p), dtype=complex)
    k0 = k
    Alpha, Beta = [], []
    beta = 0.0
    d = np.zeros(0)
    c = np.zeros(0)
    justRestarted = False
    sizeV = 1
    kk = 0
    betac = 0.0
    for mm in range(1, maxiter + 1):  # Lanczos.m:44
        for jj in range(sizeV, p + 1):  # Lanczos.m:47
            Q[:, jj - 1] = v
            r = fA(Q[:, jj - 1])
            alpha = np.real(np.vdot(r, Q[:, jj - 1]))
            if jj == 1:
                pr = fBinv(r)
                r = pr - alpha * Q[:, jj - 1]
            elif justRestarted:
                pr = fBinv(r)
                r = _gram_schmidt_gep(Q, pr, fB, jj)
                justRestarted = False
            else:
                pr = fBinv(r)
                r = pr - alpha * Q[:, jj - 1] - beta * Q[:, jj - 2]
                r = r - np.vdot(fB(Q[:, jj - 2]), r) * Q[:, jj - 2]
            r = r - np.vdot(fB(Q[:, jj - 1]), r) * Q[:, jj - 1]  # Lanczos.m:66
            r = _gram_schmidt_gep(Q, r, fB, jj)  # Lanczos.m:69
            rr = fB(r)
            beta = np.sqrt(np.vdot(r, rr))
            v = r / beta
            Alpha.append(alpha)
            Beta.append(beta)
            betac = beta
            kk += 1
        # Lanczos.m:93-98
        m = len(Alpha)
        H1 = np.diag(np.asarray(Alpha, complex)) + np.diag(np.asarray(Beta[:-1], complex), 1) \
            + np.diag(np.asarray(Beta[:-1], complex), -1)
        kd = len(d)
        H = np.zeros((kd + m, kd + m), dtype=complex)
        H[:kd, :kd] = np.diag(d)
        H[kd:, kd:] = H1
        if kd:
            H[:kd, kd] = np.conj(c)
            H[kd, :kd] = c
        Alpha, Beta = [], []
        dd, U = np.linalg.eig(H)  # Lanczos.m:103
        res = np.abs(beta * U[-1, :])
        ind = np.argsort(-np.real(dd), kind="stable")
        dd = dd[ind]
        res = res[ind]
        isNotConverged = ~(res[:k0] < tol * np.maximum(eps ** (2.0 / 3.0), np.abs(dd[:k0])))
        nconv = int(np.count_nonzero(~isNotConverged))
        if nconv >= k0 or mm == maxiter:
            break
        k = k0 + min(nconv, (p - k0) // 2)  # Lanczos.m:122
        if k == 1 and p > 3:
            k = p // 2
        ind = ind[:k]
        U = U[:, ind]
        Q[:, :k] = Q @ U
        d = dd[:k]
        c = betac * U[-1, :]
        justRestarted = True
        sizeV = k + 1
    U = U[:, ind[:k0]]
    d = dd[:k0]
    return Q @ U, d, kk


def _gram_schmidt_gep(V, r, fB, jj):
    """GramSchmitGEP (fun_Lanczos/Lanczos.m:159-167)."""
    for k in range(jj):
        r = r - np.vdot(fB(V[:, k]), r) * V[:, k]
    return r


# --------------------------------------------------------------------------
# Restoration and residuals (fun_no_inv/)
# --------------------------------------------------------------------------


def material_B(mtx_B, e, h):
    """[N e + J h; J' e + M h]   (EvaluteError_Chiral_no_inv.m:43)."""
    return mtx_B["N_int"] @ e + mtx_B["J_int"] @ h, mtx_B["J_int"].conj().T @ e + mtx_B["M_int"] @ h


def eigen_restoration(EV, fnc, mtx_B):
    """(e, h) from NFSEP eigenvectors (fun_no_inv/Eigen_Restoration_Chiral_no_inv.m:3-25).

    e_h = -i * [N Qr(v_b) + J Pr(v_a); J' Qr(v_b) + M Pr(v_a)].
    """
    EV = np.atleast_2d(EV.T).T
    es, hs = [], []
    for j in range(EV.shape[1]):
        v = EV[:, j]
        h2 = v.shape[0] // 2
        ee, hh = material_B(mtx_B, fnc.Qr(v[h2:]), fnc.Pr(v[:h2]))
        es.append(-1j * ee)
        hs.append(-1j * hh)
    return np.stack(es, 1), np.stack(hs, 1)


def evaluate_error(ev, ew, e, h, C, fA, fB, mtx_B):
    """NFSEP and original-problem residuals (fun_no_inv/EvaluteError_Chiral_no_inv.m:43-61)."""
    ev = np.atleast_2d(ev.T).T
    res_nf, res_org = [], []
    for j in range(len(ew)):
        res_nf.append(np.linalg.norm(fB(ev[:, j]) - ew[j] * fA(ev[:, j])))
        be, bh = material_B(mtx_B, C.conj().T @ h[:, j], -(C @ e[:, j]))
        res_org.append(np.linalg.norm(np.concatenate([be, bh]) - 1j * ew[j] * np.concatenate([e[:, j], h[:, j]])))
    return np.array(res_nf), np.array(res_org)


# --------------------------------------------------------------------------
# Convenience: everything for one k on one structure
# --------------------------------------------------------------------------


def vertex_material(eps_v, gam_v, mu, A):
    """Per-vertex (N-slot, M-slot, J-slot) tensors, each (N, 3, 3), of the bi-isotropic media
    eps_v I, mu I, xi = i gamma_v I, zeta = -i gamma_v I: the inverse 6x6 by a batched general
    inverse and the covariant form A' X A (Physical_settings_woodpile.m:91-96,184-194), stored
    per vertex as Material_Matrix_Generate_Mult_Point.m:38-48 stores the A{i,j} diagonals."""
    C = _bi_isotropic_6x6(eps_v, gam_v, mu)
    Ci = np.linalg.inv(C)
    return _slots_cov(Ci, A)


def vertex_material_derivative(eps_v, gam_v, mu, A, wrt):
    """d(slots)/d eps_v (wrt = 'eps') or d/d gamma_v ('gamma') per vertex: -C^-1 (dC) C^-1."""
    C = _bi_isotropic_6x6(eps_v, gam_v, mu)
    Ci = np.linalg.inv(C)
    dC = np.zeros_like(C)
    I3 = np.eye(3)
    if wrt == "eps":
        dC[:, :3, :3] = I3
    else:
        dC[:, :3, 3:] = 1j * I3
        dC[:, 3:, :3] = -1j * I3
    return _slots_cov(-Ci @ dC @ Ci, A)


def _bi_isotropic_6x6(eps_v, gam_v, mu):
    eps_v = np.asarray(eps_v, float)
    gam_v = np.broadcast_to(np.asarray(gam_v, float), eps_v.shape)
    n = eps_v.size
    C = np.zeros((n, 6, 6), complex)
    I3 = np.eye(3)
    C[:, :3, :3] = eps_v[:, None, None] * I3
    C[:, 3:, 3:] = mu * I3
    C[:, :3, 3:] = 1j * gam_v[:, None, None] * I3
    C[:, 3:, :3] = -1j * gam_v[:, None, None] * I3
    return C


def _slots_cov(Ci, A):
    A = np.asarray(A, float)
    cov = lambda X: np.einsum("ki,nkl,lj->nij", A, X, A)  # noqa: E731  A' X A per vertex
    return cov(Ci[:, :3, :3]), cov(Ci[:, 3:, 3:]), cov(Ci[:, :3, 3:])


def quadratic_form_derivative(dims, theta, x, dN, dM, dJ, delta=None):
    """x^H (d fncB) x for the material direction (dN, dM, dJ): fncB is linear in the vertex
    tensors, so this is fncB assembled from the derivative tensors (the reference gradient)."""
    P = Problem(dims, theta, dN, dM, dJ, delta)
    return np.vdot(x, P.fncB(x))


class Problem:
    """One (grid, k, material) instance: mtx_A, fnc, mtx_B and the operator handles."""

    def __init__(self, dims, theta, N_vert, M_vert, J_vert, delta=None):
        dims = tuple(int(d) for d in dims)
        self.dims = dims
        self.N = int(np.prod(dims))
        self.theta = np.asarray(theta, float)
        self.delta = np.array([1.0 / d for d in dims]) if delta is None else np.asarray(delta, float)
        self.mtx_A = svd_matrix_generate(dims, self.theta, self.delta)
        self.fnc = Fnc(self.mtx_A)
        self.N_vert, self.M_vert, self.J_vert = N_vert, M_vert, J_vert
        self.mtx_B = interpolate_material(dims, self.theta, N_vert, M_vert, J_vert)
        self.Sigma_r = self.mtx_A["Sigma_r"]

    def fncA(self, x):
        return fncA(x, self.Sigma_r)

    def fncB(self, x):
        return fncB(x, self.fnc, self.mtx_B)

    def fncinvB(self, x, tol=1e-12, maxit=10000):
        return pcg(self.fncB, x, tol, maxit)

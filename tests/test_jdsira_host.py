"""JD-SIRA host logic on the CPU: the driver runs over a stand-in plan whose
operators are dense matrices (torch CPU tensors), so the Rayleigh-Ritz /
Ritz-pair selection / restart / locking control flow is checked without a GPU.

Pins: the converged values are the smallest positive eigenvalues of A x = w B x
from scipy's dense eigh, and every locked vector has ||A u - w B u|| <= 10 * tol
(the driver also accepts a pair whose residual stayed in (tol, 10 tol] for five
steps, GEP_AB_Herm_JDSIRA_Locking4.m:181-187)."""
import numpy as np
import pytest
import scipy.linalg


class DensePlan:
    """Operators A = fncB (Hermitian positive definite) and B = fncA (Hermitian,
    indefinite) as dense matrices, with the plan's vector primitives."""

    def __init__(self, A, B):
        import torch

        self.torch = torch
        self.device = torch.device("cpu")
        self.nb = 1
        self.N = A.shape[0] // 4
        self.A = torch.as_tensor(A)
        self.B = torch.as_tensor(B)
        self.inner = 0

    def fncB(self, x):
        return self.A @ x

    def fncA(self, x):
        return self.B @ x

    def block_dot(self, V, w):
        return V.conj() @ w

    def block_gemv(self, V, y, out=None, beta=0.0):
        r = y.to(V.dtype) @ V
        if out is None:
            return r
        out.mul_(beta).add_(r)
        return out

    def mgs(self, V, t, coef=False):
        cs = []
        for j in range(V.shape[0]):  # Locking4.m:326-330
            c = self.torch.vdot(V[j], t)
            t -= c * V[j]
            cs.append(c)
        return self.torch.stack(cs) if coef else None

    def minres(self, x, sigma=0.0, tol=1e-2, maxit=30000):
        from oracle import fame_oracle as fo

        M = (self.A - sigma * self.B).numpy()
        y, _, relres, it = fo.minres(lambda v: M @ v, x.numpy(), tol, maxit)
        self.inner += it
        return self.torch.as_tensor(y), np.array([it]), np.array([relres])


    def jd_ritz_residual(self, V, y, theta):
        u = self.torch.as_tensor(np.asarray(y, complex)) @ V
        u = u / self.torch.linalg.vector_norm(u)
        r = self.fncB(u) - complex(theta) * self.fncA(u)
        return u, r, float(self.torch.linalg.vector_norm(r))

    def jd_expand(self, V, m, r, sigma, tol, maxit):
        t, its, _ = self.minres(r, sigma=sigma, tol=tol, maxit=maxit)
        nrm = float(self.torch.linalg.vector_norm(t))
        if nrm < 1e-9:
            t = t / nrm
        self.mgs(V[:m], t)
        V[m] = t / self.torch.linalg.vector_norm(t)
        return int(its[0]), (V[:m + 1].conj() @ self.fncB(V[m])).numpy(), (V[:m + 1].conj() @ self.fncA(V[m])).numpy()


def random_pencil(n, seed):
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(rng.uniform(1.0, 3.0, n)) @ Q.conj().T
    A = 0.5 * (A + A.conj().T)
    Q2 = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    d = rng.uniform(0.05, 1.0, n) * np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    B = Q2 @ np.diag(d) @ Q2.conj().T
    B = 0.5 * (B + B.conj().T)
    return A, B


def smallest_positive(A, B, k):
    mu = scipy.linalg.eigh(B, A, eigvals_only=True)  # B y = mu A y, w = 1/mu
    w = 1.0 / mu[mu > 0]
    return np.sort(w)[:k]


@pytest.mark.parametrize("seed", [0, 1])
def test_jdsira_dense_pencil(pkg, seed):
    A, B = random_pencil(64, seed)
    plan = DensePlan(A, B)
    k = 4
    res = pkg.jdsira.GEP_AB_Herm_JDSIRA_Driver4(plan, k, stop_tolerance=1e-8, no_restart=20, seed=seed)
    want = smallest_positive(A, B, k)
    got = np.sort(res.ew.real)
    assert np.allclose(got, want, rtol=1e-10), (got, want)
    assert np.all(np.abs(res.ew.imag) < 1e-10)
    EV = res.ev.numpy()
    for j in range(k):
        u = EV[j]
        assert abs(np.linalg.norm(u) - 1) < 1e-12
        assert np.linalg.norm(A @ u - res.ew[j] * (B @ u)) <= 1e-7
    assert res.total_inner > 0


def test_jdsira_shift_invert_inner(pkg):
    """A shifted inner operator A - sigma B changes the path, not the converged pairs."""
    A, B = random_pencil(48, 3)
    plan = DensePlan(A, B)
    res = pkg.jdsira.GEP_AB_Herm_JDSIRA_Driver4(plan, 3, no_restart=20, sigma_inner=0.5)
    assert np.allclose(np.sort(res.ew.real), smallest_positive(A, B, 3), rtol=1e-10)


def test_gram_schmidt_drops_dependent(pkg):
    import torch

    rng = np.random.default_rng(5)
    Y = rng.standard_normal((5, 30)) + 1j * rng.standard_normal((5, 30))
    Y[2] = 2 * Y[0] - 1j * Y[1]          # dependent: dropped
    Y[4] = 0                             # zero: dropped
    T = torch.as_tensor(Y.copy())
    plan = DensePlan(np.eye(32), np.eye(32))
    n = pkg.jdsira.gram_schmidt(plan, T)
    assert n == 3
    Q = T[:n].numpy()
    assert np.allclose(Q.conj() @ Q.T, np.eye(3), atol=1e-12)
    # span is preserved
    P = Q.T @ Q.conj()
    for i in (0, 1, 3):
        assert np.linalg.norm(P @ Y[i] - Y[i]) < 1e-10 * np.linalg.norm(Y[i])


def test_choose_ritz_pair_closest_right_of_target(pkg):
    ew = np.array([-0.5, 0.9, 0.2, 3.0, 0.4 + 1e-3j, 1e-6 + 0.5j])
    VRR = np.eye(6, dtype=complex)
    ewno, rew, rev = pkg.jdsira.choose_ritz_pair(6, ew, 1, 0.0, VRR, 1e-8, "CL", 4, 0.0, "CL",
                                                 np.zeros(8, complex), 0)
    # 'CL': real part right of the target and (almost) real, or a near-imaginary value
    assert ewno == 4
    assert np.allclose(rew[:4], [1e-6 + 0.5j, 0.2, 0.9, 3.0])
    assert np.allclose(np.abs(rev[:, 0]), VRR[:, 5])


def test_choose_ritz_pair_skips_converged(pkg):
    ew = np.array([0.2, 0.3, 0.7, 1.1])
    VRR = np.eye(4, dtype=complex)
    zf = np.zeros(8, complex)
    zf[0] = 0.2
    ewno, rew, _ = pkg.jdsira.choose_ritz_pair(4, ew, 2, 0.0, VRR, 1e-8, "CL", 4, 0.0, "CL", zf, 1)
    assert rew[0] == 0.3 and 0.2 not in rew[:ewno - 1]

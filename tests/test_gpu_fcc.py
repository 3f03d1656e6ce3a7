"""GPU parity of the FCC skewed-grid operators (C-ABI fame_fcc_*) against the FCC oracle
(oracle/fcc_oracle.py, itself pinned by unitarity, the closed form and Gamma_d T = T Lambda_d).

Tolerance: 1e-12 relative (FP64 FFT-based operators)."""
import numpy as np
import pytest

from oracle import fcc_oracle as fc

pytestmark = pytest.mark.gpu
RTOL = 1e-12
rng = np.random.default_rng(31)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def make(pkg, dims, k):
    n1, n2, n3 = dims
    lat = fc.fcc_lattice()
    delta = np.array([1 / n1, np.sqrt(3) / 2 / n2, np.sqrt(6) / 3 / n3])
    tab = fc.fft_tables(n1, n2, n3, k, *lat, delta)
    return pkg.FccPlan(dims, k, lat, delta), tab


K = np.array([0.31, -0.17, 0.23])
DIMS = [(2, 3, 3), (4, 6, 6), (6, 3, 4), (12, 12, 12), (10, 9, 8), (24, 24, 24)]
FAST = [8, 12, 16, 20, 24, 32, 40, 48, 60, 64, 72, 80, 90, 96]


@pytest.mark.parametrize("dims", DIMS)
def test_T_TH_match_oracle(pkg, torch_mod, dims):
    say(f"FCC T/TH dims={dims}")
    plan, tab = make(pkg, dims, K)
    n = plan.n
    x = crand(3, n)
    y = plan.T(torch_mod.as_tensor(x)).cpu().numpy()
    z = plan.TH(torch_mod.as_tensor(x)).cpu().numpy()
    for r in range(3):
        assert rel(y[r], fc.T(x[r], tab)) < RTOL
        assert rel(z[r], fc.TH(x[r], tab)) < RTOL
    plan.close()


def test_T_batches_beyond_one_launch_and_k0(pkg, torch_mod):
    say("FCC T nvec=50, k=0")
    dims = (4, 6, 6)
    plan, tab = make(pkg, dims, np.zeros(3))
    x = crand(50, plan.n)
    y = plan.T(torch_mod.as_tensor(x)).cpu().numpy()
    assert max(rel(y[r], fc.T(x[r], tab)) for r in range(50)) < RTOL
    plan.close()


def test_T_unitary_n60(pkg, torch_mod):
    say("FCC T unitary n=60")
    dims = (60, 60, 60)
    plan, tab = make(pkg, dims, K)
    x = torch_mod.as_tensor(crand(2, plan.n), device="cuda")
    y = plan.T(x)
    assert abs(float(y.norm() / x.norm()) - 1) < 1e-13
    assert float((plan.TH(y) - x).norm() / x.norm()) < 1e-13
    xr = x[0].cpu().numpy()
    assert rel(y[0].cpu().numpy(), fc.T(xr, tab)) < RTOL
    plan.close()


@pytest.mark.parametrize("dims", [(4, 6, 6), (6, 3, 4)])
def test_sigma_and_lambda_maps(pkg, torch_mod, dims):
    say(f"FCC Sigma_r / Pr / Qr / Prs / Qrs dims={dims}")
    plan, tab = make(pkg, dims, K)
    svd = fc.svd_curl(tab)
    n = plan.n
    assert rel(plan.Sigma_r().cpu().numpy(), svd["Sigma_r"]) < RTOL
    T = lambda v: np.concatenate([fc.T(v[i * n:(i + 1) * n], tab) for i in range(v.size // n)])  # noqa: E731
    TH = lambda v: np.concatenate([fc.TH(v[i * n:(i + 1) * n], tab) for i in range(v.size // n)])  # noqa: E731
    u = crand(16 * n)
    w = crand(24 * n)
    t = torch_mod
    assert rel(plan.Pr(t.as_tensor(u)).cpu().numpy(), T(svd["Lambda_Pr"] @ u)) < RTOL
    assert rel(plan.Qr(t.as_tensor(u)).cpu().numpy(), T(svd["Lambda_Qr"] @ u)) < RTOL
    assert rel(plan.Prs(t.as_tensor(w)).cpu().numpy(), svd["Lambda_Pr"].conj().T @ TH(w)) < RTOL
    assert rel(plan.Qrs(t.as_tensor(w)).cpu().numpy(), svd["Lambda_Qr"].conj().T @ TH(w)) < RTOL
    plan.close()


def test_nfgep_and_eigenvector(pkg, torch_mod):
    say("FCC NFGEP matvec / Compute_eigenvector")
    dims = (4, 6, 6)
    plan, tab = make(pkg, dims, K)
    svd = fc.svd_curl(tab)
    n = plan.n
    x = crand(32 * n)
    xi, zeta, phi = crand(24 * n), crand(24 * n), crand(24 * n) + 3
    t = torch_mod
    y = plan.mtx_prod_vec_A_NFGEP(t.as_tensor(x), t.as_tensor(xi), t.as_tensor(zeta), t.as_tensor(phi))
    assert rel(y.cpu().numpy(), fc.nfgep_A_matvec(x, tab, svd, xi, zeta, phi)) < RTOL
    y = plan.mtx_prod_vec_A_NFGEP(t.as_tensor(x), 0.3 - 0.2j, 0.3 + 0.2j, 2.5)
    assert rel(y.cpu().numpy(), fc.nfgep_A_matvec(x, tab, svd, 0.3 - 0.2j, 0.3 + 0.2j, 2.5)) < RTOL
    ev = plan.Compute_eigenvector(t.as_tensor(x), t.as_tensor(xi), t.as_tensor(zeta), t.as_tensor(phi))
    assert rel(ev.cpu().numpy(), fc.compute_eigenvector(x, tab, svd, xi, zeta, phi)) < RTOL
    plan.close()


@pytest.mark.parametrize("n", [8, 12])
def test_nfgep_fused_matches_oracle(pkg, torch_mod, n):
    """Cubic grids of a specialised length run the fused NFGEP kernels (Lambda products and
    the T z stage in one launch, T xy + mid + T^H xy in one, T^H z + adjoint products in one)."""
    say(f"FCC fused NFGEP / Compute_eigenvector n={n}")
    plan, tab = make(pkg, (n, n, n), K)
    svd = fc.svd_curl(tab)
    N = plan.n
    x = crand(32 * N)
    xi, zeta, phi = crand(24 * N), crand(24 * N), crand(24 * N) + 3
    t = torch_mod
    y = plan.mtx_prod_vec_A_NFGEP(t.as_tensor(x), t.as_tensor(xi), t.as_tensor(zeta), t.as_tensor(phi))
    assert rel(y.cpu().numpy(), fc.nfgep_A_matvec(x, tab, svd, xi, zeta, phi)) < RTOL
    ev = plan.Compute_eigenvector(t.as_tensor(x), t.as_tensor(xi), t.as_tensor(zeta), t.as_tensor(phi))
    assert rel(ev.cpu().numpy(), fc.compute_eigenvector(x, tab, svd, xi, zeta, phi)) < RTOL
    plan.close()


@pytest.mark.parametrize("n", FAST)
def test_nfgep_specialised_equals_generic(pkg, torch_mod, n):
    """The specialised NFGEP path (fused kernels up to length 64, separate specialised passes
    above) gives the runtime-radix path's matvec and eigenvector map to rounding."""
    say(f"FCC specialised NFGEP == generic n={n}")
    plan, _ = make(pkg, (n, n, n), K)
    N = plan.n
    t = torch_mod
    dev = dict(device="cuda")
    x = t.as_tensor(crand(32 * N), **dev)
    xi, zeta = t.as_tensor(crand(24 * N), **dev), t.as_tensor(crand(24 * N), **dev)
    phi = t.as_tensor(crand(24 * N) + 3, **dev)
    yf = plan.mtx_prod_vec_A_NFGEP(x, xi, zeta, phi)
    ef = plan.Compute_eigenvector(x, xi, zeta, phi)
    plan.set_generic(True)
    yg = plan.mtx_prod_vec_A_NFGEP(x, xi, zeta, phi)
    eg = plan.Compute_eigenvector(x, xi, zeta, phi)
    assert float((yf - yg).norm() / yg.norm()) < 1e-13
    assert float((ef - eg).norm() / eg.norm()) < 1e-13
    plan.close()


def test_material_operators(pkg, torch_mod):
    say("FCC mtx_Eps_Mu_Zeta_Xi / GEP B")
    dims = (4, 6, 6)
    plan, _ = make(pkg, dims, K)
    n = plan.n
    edge = [rng.random(n) < 0.4 for _ in range(4)]
    face = [np.nonzero(rng.random(n) < 0.3)[0] for _ in range(4)]     # index arrays
    face_b = [np.isin(np.arange(n), f) for f in face]
    tin = [crand(3, 3) for _ in range(4)]
    tair = [crand(3, 3) for _ in range(4)]
    plan.mtx_Eps_Mu_Zeta_Xi(edge, face, tin, tair)
    m = fc.eps_mu_zeta_xi(n, edge, face_b, tin, tair)
    t = torch_mod
    x = crand(12 * n)
    for name in pkg.FccPlan.PARAMS:
        assert rel(plan.param_apply(name, t.as_tensor(x)).cpu().numpy(), m[name] @ x) < RTOL, name
    x = crand(48 * n)
    assert rel(plan.GEP_B(t.as_tensor(x)).cpu().numpy(), fc.gep_B(m) @ x) < RTOL
    plan.close()


@pytest.mark.parametrize("n", FAST)
def test_specialised_T_equals_generic(pkg, torch_mod, n):
    """The compile-time four-step kernels (n1 = n2 = n3 with a specialised length) give the
    runtime-radix kernels' T and T^H to rounding."""
    say(f"FCC specialised == generic n={n}")
    plan, _ = make(pkg, (n, n, n), K)
    x = torch_mod.as_tensor(crand(2, plan.n), device="cuda")
    yf, zf = plan.T(x), plan.TH(x)
    plan.set_generic(True)
    yg, zg = plan.T(x), plan.TH(x)
    assert float((yf - yg).norm() / yg.norm()) < 1e-13
    assert float((zf - zg).norm() / zg.norm()) < 1e-13
    plan.close()


# ---------------------------------------------------------------------------
# NFGEP with the mtx_data that GEP B fixes (half-coupled xi_d / zeta_d): the oracle's
# maps are pinned against the explicit (GEP A, GEP B) in tests/test_fcc_oracle_pins.py
# ---------------------------------------------------------------------------

def physical_material(pkg, plan, n, gamma=0.7, eps_in=(12.0, 13.0, 14.0), seed=7):
    r = np.random.default_rng(seed)
    edge = [r.random(n) < 0.4 for _ in range(4)]
    face = [r.random(n) < 0.4 for _ in range(4)]
    I3 = np.eye(3)
    tin = [np.diag(eps_in).astype(complex), I3, -1j * gamma * I3, 1j * gamma * I3]
    tair = [I3, I3, 0 * I3, 0 * I3]
    plan.mtx_Eps_Mu_Zeta_Xi(edge, face, tin, tair)
    return fc.eps_mu_zeta_xi(n, edge, face, tin, tair)


@pytest.mark.parametrize("dims", [(2, 3, 3), (4, 6, 6), (8, 8, 8), (12, 12, 12)])
def test_nfgep_physical_data_matches_oracle(pkg, torch_mod, dims):
    """fame_fcc_nfgep_data = the oracle's nfgep_data; the half-coupled NFGEP matvec and
    eigenvector map (generic kernels, and the fused ones at 8^3 and 12^3) = the oracle's."""
    say(f"FCC NFGEP with GEP-B mtx_data dims={dims}")
    plan, tab = make(pkg, dims, K)
    svd = fc.svd_curl(tab)
    n = plan.n
    m = physical_material(pkg, plan, n)
    xi, zeta, phi = (v.cpu().numpy() for v in plan.nfgep_data())
    rx, rz, rp = fc.nfgep_data(m)
    assert rel(xi, rx) < 1e-15 and rel(zeta, rz) < 1e-15 and rel(phi, rp) < 1e-15
    t = torch_mod
    x = crand(32 * n)
    args = [t.as_tensor(v) for v in (xi, zeta, phi)]
    y = plan.mtx_prod_vec_A_NFGEP(t.as_tensor(x), *args, coupling=1)
    assert rel(y.cpu().numpy(), fc.nfgep_A_matvec(x, tab, svd, rx, rz, rp, coupling=1)) < RTOL
    ev = plan.Compute_eigenvector(t.as_tensor(x), *args, coupling=1)
    assert rel(ev.cpu().numpy(), fc.compute_eigenvector(x, tab, svd, rx, rz, rp, coupling=1)) < RTOL
    plan.close()


def test_nfgep_eigenvector_solves_gep_on_device(pkg, torch_mod):
    """NFGEP eigenpairs from the device matvec (dense, 2x3x3): their eigenvalues are the nonzero
    eigenvalues of the explicit (GEP A, GEP B), and the device Compute_eigenvector, mapped by
    diag(I, iI), solves GEP_A x = lambda GEP_B x."""
    import scipy.linalg as sl
    import scipy.sparse as sp

    say("FCC NFGEP eigenpairs on the device vs the explicit GEP, 2x3x3")
    dims = (2, 3, 3)
    plan, tab = make(pkg, dims, K)
    svd = fc.svd_curl(tab)
    n = plan.n
    m = physical_material(pkg, plan, n, gamma=1.1)
    data = plan.nfgep_data()
    t = torch_mod
    E = t.eye(32 * n, dtype=t.complex128, device="cuda")
    NF = np.stack([plan.mtx_prod_vec_A_NFGEP(E[j], *data, coupling=1).cpu().numpy() for j in range(32 * n)], 1)
    AF = np.stack([fc.fncA(e, svd) for e in np.eye(32 * n, dtype=complex)], 1)
    w, Z = sl.eig(NF, AF)
    a1, a2, a3 = fc.fcc_lattice()
    delta = np.array([1 / 2, np.sqrt(3) / 2 / 3, np.sqrt(6) / 3 / 3])
    blocks, _ = fc.curl_blocks(*dims, K, a1, a2, a3, delta)
    A, _ = fc.gep_A(blocks)
    B = fc.gep_B(m)
    wg = sl.eigvals(A.toarray(), B.toarray())
    wg = wg[np.abs(wg) > 1e-9 * np.abs(wg).max()]
    assert max(np.min(np.abs(wg - x)) / abs(x) for x in w) < 1e-10
    P = sp.block_diag([sp.identity(24 * n), 1j * sp.identity(24 * n)]).tocsr()
    for i in np.argsort(np.abs(w))[:6]:
        ev = plan.Compute_eigenvector(t.as_tensor(Z[:, i]), *data, coupling=1).cpu().numpy()
        x = P @ ev
        res = np.linalg.norm(A @ x - w[i] * (B @ x)) / (np.linalg.norm(x) * abs(w[i]))
        assert res < 1e-11, (w[i], res)
    plan.close()


@pytest.mark.parametrize("n", [8, 24, 48, 60, 64])
def test_nfgep_coupled_specialised_equals_generic(pkg, torch_mod, n):
    """The half-coupled NFGEP on the fused path (pairing vector u with 24 + (u + 12) mod 24) equals
    the generic path to rounding."""
    say(f"FCC coupled NFGEP specialised == generic n={n}")
    plan, _ = make(pkg, (n, n, n), K)
    N = plan.n
    t = torch_mod
    dev = dict(device="cuda")
    x = t.as_tensor(crand(32 * N), **dev)
    xi, zeta = t.as_tensor(crand(24 * N), **dev), t.as_tensor(crand(24 * N), **dev)
    phi = t.as_tensor(crand(24 * N) + 3, **dev)
    yf = plan.mtx_prod_vec_A_NFGEP(x, xi, zeta, phi, coupling=1)
    ef = plan.Compute_eigenvector(x, xi, zeta, phi, coupling=1)
    plan.set_generic(True)
    yg = plan.mtx_prod_vec_A_NFGEP(x, xi, zeta, phi, coupling=1)
    eg = plan.Compute_eigenvector(x, xi, zeta, phi, coupling=1)
    assert float((yf - yg).norm() / yg.norm()) < 1e-13
    assert float((ef - eg).norm() / eg.norm()) < 1e-13
    plan.close()


def test_svd_maps_refuse_gamma_point(pkg, torch_mod):
    """At k = 0 the FCC curl SVD divides by zero (mode (0,0,0)): T and T^H still run, the SVD
    maps raise instead of returning Inf/NaN."""
    say("FCC k=0: SVD maps refused")
    plan, _ = make(pkg, (4, 6, 6), np.zeros(3))
    n = plan.n
    t = torch_mod
    for call in (lambda: plan.Sigma_r(), lambda: plan.Pr(t.as_tensor(crand(16 * n))),
                 lambda: plan.mtx_prod_vec_A_NFGEP(t.as_tensor(crand(32 * n)), 0.1, 0.1, 2.0)):
        with pytest.raises(pkg.FameError, match="Gamma"):
            call()
    plan.close()


@pytest.mark.parametrize("n", [8, 24, 48, 60, 64])
@pytest.mark.parametrize("coupling", [0, 1])
def test_nfgep_split_xy_kernel_equals_paired(pkg, torch_mod, n, coupling):
    """The fused NFGEP with its xy kernel transforming one plane at a time (one plane of LDS)
    gives the two-planes-in-LDS kernel's matvec and eigenvector map, and the generic path's."""
    say(f"FCC NFGEP split xy kernel n={n} coupling={coupling}")
    plan, _ = make(pkg, (n, n, n), K)
    N = plan.n
    t = torch_mod
    dev = dict(device="cuda")
    x = t.as_tensor(crand(32 * N), **dev)
    xi, zeta = t.as_tensor(crand(24 * N), **dev), t.as_tensor(crand(24 * N), **dev)
    phi = t.as_tensor(crand(24 * N) + 3, **dev)
    outs = []
    for split in (0, 1):
        plan.set_nfgep_variant(split)
        outs.append((plan.mtx_prod_vec_A_NFGEP(x, xi, zeta, phi, coupling=coupling),
                     plan.Compute_eigenvector(x, xi, zeta, phi, coupling=coupling)))
    plan.set_generic(True)
    yg = plan.mtx_prod_vec_A_NFGEP(x, xi, zeta, phi, coupling=coupling)
    eg = plan.Compute_eigenvector(x, xi, zeta, phi, coupling=coupling)
    for y, e in outs:
        assert float((y - yg).norm() / yg.norm()) < 1e-13
        assert float((e - eg).norm() / eg.norm()) < 1e-13
    plan.close()

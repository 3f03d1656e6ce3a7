"""GPU parity of the per-vertex material path and the adjoint gradient kernel.

* fncB with a per-vertex bi-isotropic material equals the oracle's fncB assembled from
  the same per-vertex tensors (1e-12), and equals the id-table path when the vertex
  values repeat a table's two media (1e-13);
* fame_material_gradient: the vertex terms sum to x^H fncB x; directional derivatives
  match the reference gradient assembled from the oracle's sparse matrices with the
  derivative tensors (fncB is linear in the vertex tensors), to 1e-10;
* eigenvalue gradients dw/d eps_v, dw/d gamma match central finite differences
  (step 1e-6) of dense eigenvalues to 1e-6 relative.
"""
import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(505)


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


def pv_setup(pkg, dims, gamma=0.2, seed=5):
    S = pkg.structures
    ids = S.woodpile_sc_ids(dims)
    r = np.random.default_rng(seed)
    eps = np.where(ids == 1, r.uniform(1, 13, ids.size), 1.0)
    gam = np.where(ids == 1, gamma, 0.0)
    return ids, eps, gam


THETAS = np.array([[0.4, -1.3, 2.2], [1.5, 0.1, -0.7]])


# a skewed lattice (non-diagonal metric G = A^T A) and mu != 1 exercise every slot of the
# closed-form per-vertex product; "generic" runs the two-kernel form instead of the z-march
SKEW = np.array([[1.0, 0.5, 0.2], [0.0, 0.9, 0.3], [0.1, 0.0, 1.1]])


@pytest.mark.parametrize("dims,lattice,mu,generic", [
    ((6, 5, 4), "eye", 1.0, False), ((8, 8, 8), "eye", 1.0, False),
    ((8, 8, 8), "skew", 1.7, False), ((6, 5, 4), "skew", 1.7, False), ((8, 8, 8), "skew", 1.7, True)])
def test_vertex_material_fncB_matches_oracle(pkg, fo, torch_mod, dims, lattice, mu, generic):
    say(f"per-vertex fncB dims={dims} lattice={lattice} mu={mu} generic={generic}")
    A = np.eye(3) if lattice == "eye" else SKEW
    ids, eps, gam = pv_setup(pkg, dims)
    table = pkg.structures.bi_isotropic_table(13.0, 0.2, np.eye(3))
    plan = pkg.Plan(dims, THETAS, table, ids)
    plan.set_vertex_material(eps, gam, mu, A)
    if generic:
        plan.set_generic(True)
    x = crand(2, 4 * plan.N)
    y = plan.fncB(torch_mod.as_tensor(x)).cpu().numpy()
    Nv, Mv, Jv = fo.vertex_material(eps, gam, mu, A)
    for b in range(2):
        P = fo.Problem(dims, THETAS[b], Nv, Mv, Jv)
        assert rel(y[b], P.fncB(x[b])) < 1e-12
    plan.close()


def test_vertex_material_equals_table(pkg, torch_mod):
    say("per-vertex == id table")
    dims = (12, 12, 12)
    S = pkg.structures
    ids = S.woodpile_sc_ids(dims)
    table = S.bi_isotropic_table(13.0, 0.2, np.eye(3))
    pt = pkg.Plan(dims, THETAS, table, ids)
    pv = pkg.Plan(dims, THETAS, table, ids)
    pv.set_vertex_material(np.where(ids == 1, 13.0, 1.0), np.where(ids == 1, 0.2, 0.0))
    x = torch_mod.as_tensor(crand(2, 4 * pt.N))
    assert rel(pv.fncB(x).cpu().numpy(), pt.fncB(x).cpu().numpy()) < 1e-13
    pt.close()
    pv.close()


@pytest.mark.parametrize("dims", [(6, 5, 4), (8, 8, 8)])
def test_material_gradient_matches_reference(pkg, fo, torch_mod, dims):
    say(f"material gradient vs assembled reference dims={dims}")
    ids, eps, gam = pv_setup(pkg, dims, seed=7)
    theta = THETAS[0]
    plan = pkg.Plan(dims, theta, pkg.structures.bi_isotropic_table(13.0, 0.2, np.eye(3)), ids)
    plan.set_vertex_material(eps, gam)
    N = plan.N
    x = crand(3, 4 * N)
    g = plan.material_gradient(torch_mod.as_tensor(x)).cpu().numpy()   # (3, 3, N)
    Nv, Mv, Jv = fo.vertex_material(eps, gam, 1.0, np.eye(3))
    P = fo.Problem(dims, theta, Nv, Mv, Jv)
    for j in range(3):
        q = np.vdot(x[j], P.fncB(x[j])).real
        assert abs(g[j, 2].sum() - q) <= 1e-12 * abs(q)
    for wrt, row in (("eps", 0), ("gamma", 1)):
        dN, dM, dJ = fo.vertex_material_derivative(eps, gam, 1.0, np.eye(3), wrt)
        for trial in range(2):
            dirn = rng.standard_normal(N)
            ref = fo.quadratic_form_derivative(dims, theta, x[trial], dN * dirn[:, None, None],
                                               dM * dirn[:, None, None], dJ * dirn[:, None, None])
            got = g[trial, row] @ dirn
            assert abs(ref.imag) < 1e-10 * abs(ref)
            assert abs(got - ref.real) <= 1e-10 * abs(ref.real), (wrt, trial)
        for v in rng.choice(N, 3, replace=False):  # single vertices
            e1 = np.zeros(N)
            e1[v] = 1
            ref = fo.quadratic_form_derivative(dims, theta, x[2], dN * e1[:, None, None], dM * e1[:, None, None],
                                               dJ * e1[:, None, None]).real
            assert abs(g[2, row, v] - ref) <= 1e-10 * max(abs(ref), 1e-300) + 1e-14 * np.abs(g[2, row]).max()
    plan.close()


def _dense_pencil(fo, dims, theta, eps, gam):
    Nv, Mv, Jv = fo.vertex_material(eps, gam, 1.0, np.eye(3))
    P = fo.Problem(dims, theta, Nv, Mv, Jv)
    n4 = 4 * P.N
    I = np.eye(n4, dtype=complex)
    Ad = np.stack([P.fncA(e) for e in I], 1)
    Bd = np.stack([P.fncB(e) for e in I], 1)
    return Ad, Bd


def _lowest(Ad, Bd, k):
    mu, X = scipy.linalg.eigh(Ad, Bd)          # Ad x = mu Bd x, omega = 1/mu
    pos = np.nonzero(mu > 0)[0]
    order = pos[np.argsort(1.0 / mu[pos])][:k]
    return 1.0 / mu[order], X[:, order]


def test_eigen_gradient_vs_finite_differences(pkg, fo, torch_mod):
    say("eigenvalue gradients vs central FD (step 1e-6), n=4")
    dims = (4, 4, 4)
    S = pkg.structures
    ids, eps, gam = pv_setup(pkg, dims, seed=11)
    eps = np.random.default_rng(12).uniform(1, 13, ids.size)          # every voxel random
    theta = S.theta_of_k(np.pi * np.array([0.5, 0.3, 0.1]), S.SC_LATTICE)
    plan = pkg.Plan(dims, theta, S.bi_isotropic_table(13.0, 0.2, np.eye(3)), ids)
    plan.set_vertex_material(eps, gam)
    k = 5
    Ad, Bd = _dense_pencil(fo, dims, theta, eps, gam)
    w0, X = _lowest(Ad, Bd, k)
    assert np.all(np.diff(w0) > 1e-6 * w0[:-1])      # simple eigenvalues at this k
    ev = torch_mod.as_tensor(X.T.copy())[None]
    d_eps, d_gam = pkg.adjoint.eigen_gradients(plan, ev, w0[None])
    d_eps, d_gam = d_eps.cpu().numpy()[0], d_gam.cpu().numpy()[0]
    h = 1e-6
    for v in (3, 17, 40):
        ep, em = eps.copy(), eps.copy()
        ep[v] += h
        em[v] -= h
        wp, _ = _lowest(*_dense_pencil(fo, dims, theta, ep, gam), k)
        wm, _ = _lowest(*_dense_pencil(fo, dims, theta, em, gam), k)
        fd = (wp - wm) / (2 * h)
        assert np.allclose(d_eps[:, v], fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max()), (v, d_eps[:, v], fd)
    # gamma of every rod vertex moved together (the structure's chirality)
    rod = ids == 1
    gp, gm = gam + h * rod, gam - h * rod
    wp, _ = _lowest(*_dense_pencil(fo, dims, theta, eps, gp), k)
    wm, _ = _lowest(*_dense_pencil(fo, dims, theta, eps, gm), k)
    fd = (wp - wm) / (2 * h)
    adj = d_gam[:, rod].sum(axis=1)
    assert np.allclose(adj, fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max())
    plan.close()


def test_band_gradient_forward_backward(pkg, torch_mod):
    say("BandGradient forward/backward on the device Lanczos")
    dims = (8, 8, 8)
    S = pkg.structures
    ids, eps, gam = pv_setup(pkg, dims, seed=21)
    thetas = np.stack([S.theta_of_k(np.pi * np.array(k), S.SC_LATTICE) for k in ([0.5, 0.3, 0.1], [0.2, 0.4, 0.6])])
    plan = pkg.Plan(dims, thetas, S.bi_isotropic_table(13.0, 0.2, np.eye(3)), ids)
    plan.set_vertex_material(eps, gam)
    bg = pkg.adjoint.BandGradient(plan)
    omega = bg.forward(4, tol=1e-12, maxiter=500)
    assert omega.shape == (2, 4) and np.all(np.diff(omega, axis=1) >= 0)
    w = np.zeros((2, 4))
    w[:, 0] = 1.0
    g_eps, g_gam = bg.backward(w)
    # directional check through a second forward solve: omega(eps + t d) - omega(eps - t d)
    d = np.random.default_rng(3).standard_normal(plan.N)
    t = 1e-6
    pred = (g_eps.cpu().numpy() @ d)
    outs = []
    for sgn in (1, -1):
        plan.set_vertex_material(eps + sgn * t * d, gam)
        outs.append(pkg.adjoint.BandGradient(plan).forward(4, tol=1e-12, maxiter=500)[:, 0])
    fd = (outs[0] - outs[1]) / (2 * t)
    assert np.allclose(pred, fd, rtol=1e-5), (pred, fd)
    plan.close()

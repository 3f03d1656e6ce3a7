"""GPU acceptance tests on the benchmark configurations' own structures, through the C-ABI.

* Eigenvectors: Lanczos (fun_Lanczos/Lanczos.m) and JD-SIRA (GEP_AB_Herm_JDSIRA_Driver4.m)
  eigenvectors span the dense solution's eigenspaces; the subspace angle against dense
  ``eigh`` of the oracle's pencil is bounded by residual / gap (Davis-Kahan), including
  at k-points with exactly degenerate omega (doublets along M-R, quadruplets at R).
* Gradients: the cluster-mean derivative at a degenerate k-point and the band-4/5 gap
  derivative against central finite differences (step 1e-6) of dense eigenvalues, 1e-6
  relative above the differences' own rounding floor (FD_FLOOR: 10 eps |omega| / h).
* Config 1 (SC woodpile n = 24, one k, 10 pairs): device Lanczos against the committed
  fixture tests/golden/config1_n24_omega.json (oracle Lanczos = ARPACK to 1e-10), 1e-10.
* Config 3 (FCC-lattice spheres r = 0.12a, gamma ~ 2.72): 12-pair Lanczos at 8^3 against
  dense eigenvalues (1e-10); grids of uneven n batched = per-k plain; n = 48 / 72 invariants.
* Config 4 (near the Pasteur limit, JD-SIRA with shift-invert MINRES, target 0.3, 'CL'):
  FCC spheres gamma = 3.46427 and tetrahedra gamma = 1.009 at 8^3 against the dense pencil
  (omega 1e-10, residual <= 1e-8, subspace angles); n = 60 property checks.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
rng = np.random.default_rng(2024)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def dense_pencil(P):
    """Dense fncA, fncB of an oracle Problem."""
    n4 = 4 * P.N
    I = np.eye(n4, dtype=complex)
    return np.stack([P.fncA(e) for e in I], 1), np.stack([P.fncB(e) for e in I], 1)


def dense_omega(Ad, Bd):
    """All omega (fncB x = omega fncA x) with fncB-orthonormal eigenvectors, ascending."""
    mu, X = sl.eigh(Ad, Bd)          # fncA x = mu fncB x, omega = 1/mu
    om = 1.0 / mu
    o = np.argsort(om)
    return om[o], X[:, o]


def clusters(w, rtol=1e-8):
    groups, cur = [], [0]
    for j in range(1, len(w)):
        if abs(w[j] - w[cur[0]]) <= rtol * abs(w[cur[0]]):
            cur.append(j)
        else:
            groups.append(cur)
            cur = [j]
    groups.append(cur)
    return groups


def check_eigenspaces(omega, U, w_all, X_all, Ad, Bd, rtol=1e-10):
    """Each returned omega matches a dense one; the returned vectors of each cluster lie in the
    dense eigenspace of that cluster: sin(angle) <= 10 * residual / gap + 1e-10, where the
    residual is ||fncB u - omega fncA u|| in the fncB^{-1} norm relative to ||u||_fncB and the
    gap is the relative distance to the rest of the spectrum (Davis-Kahan for the definite
    pencil).  Returns the largest angle."""
    worst = 0.0
    Binv = np.linalg.inv(Bd)
    for g in clusters(np.asarray(omega)):
        om = np.mean(np.asarray(omega)[g])
        j = np.argmin(np.abs(w_all - om))
        assert abs(w_all[j] - om) <= rtol * abs(om), (om, w_all[j])
        members = [i for i in range(len(w_all)) if abs(w_all[i] - w_all[j]) <= 1e-8 * abs(w_all[j])]
        Ug = U[:, g]
        res = 0.0
        for c in range(Ug.shape[1]):
            u = Ug[:, c]
            r = Bd @ u - np.asarray(omega)[g[c]] * (Ad @ u)
            res = max(res, np.sqrt(abs(np.vdot(r, Binv @ r))) / np.sqrt(abs(np.vdot(u, Bd @ u))))
        rest = np.delete(np.arange(len(w_all)), members)
        gap = np.min(np.abs(1.0 / w_all[rest] - 1.0 / om)) * abs(om)   # relative gap in mu = 1/omega
        ang = float(np.max(sl.subspace_angles(Ug, X_all[:, members])))
        assert np.sin(ang) <= 10 * res / gap + 1e-10, (om, ang, res, gap)
        worst = max(worst, ang)
    return worst


# ---------------------------------------------------------------------------
# Lanczos eigenvectors (config 1/2 structure) by subspace angle
# ---------------------------------------------------------------------------

def woodpile(pkg, fo, dims, theta, gamma=0.2):
    S = pkg.structures
    table, ids = S.bi_isotropic_table(13.0, gamma, S.SC_LATTICE), S.woodpile_sc_ids(dims)
    T = table[ids]
    return pkg.Plan(dims, theta, table, ids), fo.Problem(dims, theta, T[:, 0], T[:, 1], T[:, 2])


@pytest.mark.parametrize("kname", ["generic", "k25_doublets", "R_quadruplets"])
def test_lanczos_eigenvectors_subspace_angle(pkg, fo, torch_mod, kname):
    S = pkg.structures
    ks = S.sc_band_path()
    k = {"generic": np.pi * np.array([0.5, 0.3, 0.1]), "k25_doublets": ks[25], "R_quadruplets": ks[30]}[kname]
    dims = (6, 6, 6)
    theta = S.theta_of_k(k, S.SC_LATTICE)
    say(f"Lanczos eigenvectors vs dense eigh, n=6, {kname}")
    plan, P = woodpile(pkg, fo, dims, theta)
    ev, ew, _ = plan.Lanczos(10, tol=1e-10, maxiter=10000, p=20, cg_tol=1e-12)
    omega = 1.0 / ew[0]
    o = np.argsort(omega)
    U = ev[0].cpu().numpy()[o].T
    Ad, Bd = dense_pencil(P)
    w_all, X_all = dense_omega(Ad, Bd)
    pos = w_all > 0
    assert np.allclose(omega[o], w_all[pos][:10], rtol=1e-10, atol=0)
    if kname != "generic":
        assert len(clusters(omega[o])) < 10       # the degeneracies this case is for
    ang = check_eigenspaces(omega[o], U, w_all, X_all, Ad, Bd)
    say(f"  largest principal angle {ang:.2e}")
    plan.close()


# ---------------------------------------------------------------------------
# Gradients: degenerate clusters and the band-4/5 gap (config 5)
# ---------------------------------------------------------------------------

def fd_floor(omega, h):
    """Rounding error of a central difference of dense eigenvalues: each omega carries about
    eps |omega| (LAPACK backward error), divided by the step: 10 eps max|omega| / h."""
    return 10 * np.finfo(float).eps * float(np.max(np.abs(omega))) / h


def pv_pencil(fo, dims, theta, eps, gam):
    Nv, Mv, Jv = fo.vertex_material(eps, gam, 1.0, np.eye(3))
    return dense_pencil(fo.Problem(dims, theta, Nv, Mv, Jv))


def lowest(Ad, Bd, k):
    w, X = dense_omega(Ad, Bd)
    pos = np.nonzero(w > 0)[0][:k]
    return w[pos], X[:, pos]


def test_cluster_mean_gradient_vs_fd(pkg, fo, torch_mod):
    """At k#25 of the SC band path (n = 6) the lowest omega come in exact doublets; the derivative
    of each doublet's mean by a vertex eps (eigen_gradients with omega) equals the central
    difference of the mean of the same two sorted dense eigenvalues."""
    say("cluster-mean eigenvalue gradient vs FD at a degenerate k, n=6")
    S = pkg.structures
    dims = (6, 6, 6)
    ids = S.woodpile_sc_ids(dims)
    eps = np.where(ids == 1, 13.0, 1.0)
    gam = np.where(ids == 1, 0.2, 0.0)
    theta = S.theta_of_k(S.sc_band_path()[25], S.SC_LATTICE)
    plan = pkg.Plan(dims, theta, S.bi_isotropic_table(13.0, 0.2, np.eye(3)), ids)
    plan.set_vertex_material(eps, gam)
    k = 6
    w0, X = lowest(*pv_pencil(fo, dims, theta, eps, gam), k)
    groups = clusters(w0)
    assert any(len(g) == 2 for g in groups)
    # eigenvectors of a degenerate cluster from the device Lanczos (any basis of the eigenspace)
    ev, ew, _ = plan.Lanczos(k, tol=1e-12, maxiter=10000, p=2 * k)
    om = 1.0 / ew[0]
    o = np.argsort(om)
    assert np.allclose(om[o], w0, rtol=1e-10)
    evs = ev[:, o].contiguous()
    d_eps, _ = pkg.adjoint.eigen_gradients(plan, evs, om[o][None])
    d_eps = d_eps.cpu().numpy()[0]
    h = 1e-6
    for v in np.nonzero(ids == 1)[0][[0, 7, 19]]:
        ep, em = eps.copy(), eps.copy()
        ep[v] += h
        em[v] -= h
        wp, _ = lowest(*pv_pencil(fo, dims, theta, ep, gam), k)
        wm, _ = lowest(*pv_pencil(fo, dims, theta, em, gam), k)
        for g in groups:
            fd = (wp[g].mean() - wm[g].mean()) / (2 * h)
            for j in g:
                assert abs(d_eps[j, v] - fd) <= 1e-6 * abs(fd) + fd_floor(w0, h), (v, g, d_eps[j, v], fd)
    plan.close()


def test_band_gap_gradient_vs_fd(pkg, fo, torch_mod):
    """Config 5's band-4/5 gap omega_5 - omega_4 (per-vertex eps ~ U[1, 13] in the rods, simple
    eigenvalues): BandGradient.backward with dL/domega = e_5 - e_4 against central differences."""
    say("band-4/5 gap gradient vs FD, n=4")
    S = pkg.structures
    dims = (4, 4, 4)
    ids = S.woodpile_sc_ids(dims)
    r = np.random.default_rng(5)
    eps = np.where(ids == 1, r.uniform(1, 13, ids.size), 1.0)
    gam = np.where(ids == 1, 0.2, 0.0)
    theta = S.theta_of_k(np.pi * np.array([0.5, 0.3, 0.1]), S.SC_LATTICE)
    plan = pkg.Plan(dims, theta, S.bi_isotropic_table(13.0, 0.2, np.eye(3)), ids)
    plan.set_vertex_material(eps, gam)
    bg = pkg.adjoint.BandGradient(plan)
    omega = bg.forward(8, tol=1e-12, maxiter=2000)
    w0, _ = lowest(*pv_pencil(fo, dims, theta, eps, gam), 8)
    assert np.allclose(omega[0], w0, rtol=1e-10)
    assert np.all(np.diff(w0) > 1e-6 * w0[:-1])
    dL = np.zeros((1, 8))
    dL[0, 4], dL[0, 3] = 1.0, -1.0
    g_eps, g_gam = bg.backward(dL)
    g_eps, g_gam = g_eps.cpu().numpy()[0], g_gam.cpu().numpy()[0]
    h = 1e-6
    for v in np.random.default_rng(55).choice(plan.N, 6, replace=False):
        ep, em = eps.copy(), eps.copy()
        ep[v] += h
        em[v] -= h
        wp, _ = lowest(*pv_pencil(fo, dims, theta, ep, gam), 8)
        wm, _ = lowest(*pv_pencil(fo, dims, theta, em, gam), 8)
        fd = ((wp[4] - wp[3]) - (wm[4] - wm[3])) / (2 * h)
        assert abs(g_eps[v] - fd) <= 1e-6 * abs(fd) + 2 * fd_floor(w0, h), (v, g_eps[v], fd)
    rod = ids == 1
    wp, _ = lowest(*pv_pencil(fo, dims, theta, eps, gam + h * rod), 8)
    wm, _ = lowest(*pv_pencil(fo, dims, theta, eps, gam - h * rod), 8)
    fd = ((wp[4] - wp[3]) - (wm[4] - wm[3])) / (2 * h)
    assert abs(g_gam[rod].sum() - fd) <= 1e-6 * abs(fd) + 2 * fd_floor(w0, h)
    plan.close()


# ---------------------------------------------------------------------------
# Config 1: the reference's CPU-runnable case against the committed fixture
# ---------------------------------------------------------------------------

def test_config1_n24_against_fixture(pkg, torch_mod):
    path = os.path.join(GOLDEN, "config1_n24_omega.json")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    ref = json.load(open(path))
    say("config 1: n=24 woodpile Lanczos, 10 pairs, against the fixture")
    S = pkg.structures
    dims = (24, 24, 24)
    theta = S.theta_of_k(np.pi * np.array([0.5, 0.3, 0.1]), S.SC_LATTICE)
    plan = pkg.Plan(dims, theta, S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE), S.woodpile_sc_ids(dims))
    ev, ew, it = plan.Lanczos(10, tol=1e-10, maxiter=10000, p=20, cg_tol=1e-12, cg_maxit=10000)
    omega = np.sort(1.0 / ew[0])
    want = np.asarray(ref["omega"])
    say(f"  {int(it[0])} Lanczos steps; max rel diff {np.max(np.abs(omega - want) / want):.2e}")
    assert np.allclose(omega, want, rtol=1e-10, atol=0)
    for j in range(10):
        u = ev[0, j].contiguous()
        r = plan.fncB(u) - float(1.0 / ew[0, j]) * plan.fncA(u)
        assert float(r.norm() / u.norm()) < 1e-8
    plan.close()


# ---------------------------------------------------------------------------
# Config 3: FCC-lattice spheres, several grids
# ---------------------------------------------------------------------------

GAMMA3 = 2.72


def fcc_sphere(pkg, fo, dims, ks, gamma=GAMMA3, radius=0.12, ids=None):
    S = pkg.structures
    A = S.FCC_LATTICE
    th = np.stack([S.theta_of_k(k, A) for k in np.atleast_2d(ks)])
    table = S.bi_isotropic_table(13.0, gamma, A)
    ids = S.sphere_ids(dims, radius, A) if ids is None else ids
    T = table[ids]
    probs = [fo.Problem(dims, t, T[:, 0], T[:, 1], T[:, 2]) for t in th]
    return pkg.Plan(dims, th, table, ids), probs, th, table, ids


def test_config3_lanczos_12_pairs_vs_dense(pkg, fo, torch_mod):
    say("config 3: FCC spheres gamma=2.72 n=8, 12-pair Lanczos vs dense, 2 k batched")
    S = pkg.structures
    ks = S.fcc_band_path(8)[[1, 5]]
    plan, probs, *_ = fcc_sphere(pkg, fo, (8, 8, 8), ks, radius=0.3)
    assert 0 < plan_ids_fraction(pkg, (8, 8, 8), 0.3) < 1
    ev, ew, _ = plan.Lanczos(12, tol=1e-10, maxiter=10000, p=24)
    for b, P in enumerate(probs):
        Ad, Bd = dense_pencil(P)
        w_all, X_all = dense_omega(Ad, Bd)
        om = 1.0 / ew[b]
        o = np.argsort(om)
        assert np.allclose(om[o], w_all[w_all > 0][:12], rtol=1e-10, atol=0)
        check_eigenspaces(om[o], ev[b].cpu().numpy()[o].T, w_all, X_all, Ad, Bd)
    plan.close()


def plan_ids_fraction(pkg, dims, radius):
    S = pkg.structures
    return float(S.sphere_ids(dims, radius, S.FCC_LATTICE).mean())


def test_config3_uneven_grids_batched_equals_plain(pkg, fo, torch_mod):
    """Grids of uneven n (8 and 12), several k-points each: each grid's batched solve returns the
    per-k plain solves' omega (1e-10); the multi-grid driver runs both grids."""
    say("config 3: uneven grids n=8/12 batched == plain")
    S = pkg.structures
    ks = S.fcc_band_path(12)
    jobs = [((8, 8, 8), ks[[1, 4, 7]]), ((12, 12, 12), ks[[2, 9]])]
    for dims, kk in jobs:
        plan, _, th, table, ids = fcc_sphere(pkg, fo, dims, kk, radius=0.2)
        _, ew_b, _ = plan.Lanczos(6, tol=1e-10, maxiter=10000, p=12)
        plan.close()
        for b in range(len(kk)):
            p1 = pkg.Plan(dims, th[b], table, ids)
            _, ew_p, _ = p1.Lanczos(6, tol=1e-10, maxiter=10000, p=12)
            p1.close()
            assert np.allclose(np.sort(1 / ew_b[b]), np.sort(1 / ew_p[0]), rtol=1e-10, atol=0)


@pytest.mark.parametrize("n", [48, 72])
def test_config3_full_size_invariants(pkg, torch_mod, n):
    """n = 48 / 72 FCC spheres: fncB Hermitian and positive on random vectors, and a 12-pair
    Lanczos solve whose pairs satisfy ||fncB x - omega fncA x|| <= 1e-8 ||x||."""
    say(f"config 3: FCC spheres n={n} invariants")
    S = pkg.structures
    A = S.FCC_LATTICE
    dims = (n, n, n)
    theta = S.theta_of_k(S.fcc_band_path(21)[3], A)
    plan = pkg.Plan(dims, theta, S.bi_isotropic_table(13.0, GAMMA3, A), S.sphere_ids(dims, 0.12, A))
    t = torch_mod
    x = t.as_tensor(crand(4 * plan.N), device="cuda")
    y = t.as_tensor(crand(4 * plan.N), device="cuda")
    bx, by = plan.fncB(x).reshape(-1), plan.fncB(y).reshape(-1)
    a, b = complex(t.vdot(y, bx)), complex(t.vdot(x, by))
    assert abs(a - np.conj(b)) <= 1e-12 * abs(a)
    q = complex(t.vdot(x, bx))
    assert q.real > 0 and abs(q.imag) < 1e-12 * q.real
    ev, ew, _ = plan.Lanczos(12, tol=1e-10, maxiter=10000, p=24)
    for j in range(12):
        u = ev[0, j].contiguous()
        r = plan.fncB(u) - float(1.0 / ew[0, j]) * plan.fncA(u)
        assert float(r.norm() / u.norm()) <= 1e-8
    plan.close()


# ---------------------------------------------------------------------------
# Config 4: JD-SIRA near the Pasteur limit
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("shape,gamma", [("sphere", 3.46427), ("tetra", 1.009)])
def test_config4_jdsira_vs_dense(pkg, fo, torch_mod, shape, gamma):
    """FCC lattice, eps = 13, 'CL' target 0.3 with the shifted inner operator (sigma_inner = 0.3,
    MINRES): the converged omega are the dense pencil's nearest to 0.3 among those above it (the
    reference's 'CL' admits only Ritz values with real part above the target,
    Choose_Ritz_Pair_Unsymm4.m:263-270), 1e-10; residuals <= 1e-8; eigenvectors in the dense
    eigenspaces."""
    say(f"config 4: JD-SIRA {shape} gamma={gamma}, n=8, target 0.3")
    S = pkg.structures
    A = S.FCC_LATTICE
    dims = (8, 8, 8)
    ids = S.sphere_ids(dims, 0.3, A) if shape == "sphere" else S.tetra_ids(dims, 0.6, A)
    assert 0 < ids.mean() < 1
    k = S.fcc_band_path(8)[1]
    plan, (P,), *_ = fcc_sphere(pkg, fo, dims, k, gamma=gamma, ids=ids)
    Ad, Bd = dense_pencil(P)
    w_all, X_all = dense_omega(Ad, Bd)
    nev = 6
    above = np.sort(w_all[w_all > 0.3])
    assert above[nev] - above[nev - 1] > 1e-6 * above[nev]   # the wanted set is well separated
    res = pkg.jdsira.GEP_AB_Herm_JDSIRA_Driver4(plan, nev, stop_tolerance=1e-9, target=0.3, target_type="CL",
                                                 sigma_inner=0.3, seed=0)
    got = np.sort(res.ew.real)
    want = above[:nev]
    say(f"  omega {got}")
    assert np.allclose(got, want, rtol=1e-10, atol=0)
    EV = res.ev.cpu().numpy()
    o = np.argsort(res.ew.real)
    for j in range(nev):
        u = EV[j]
        assert np.linalg.norm(Bd @ u - res.ew[j] * (Ad @ u)) <= 1e-8 * np.linalg.norm(u)
    check_eigenspaces(got, EV[o].T, w_all, X_all, Ad, Bd)
    plan.close()


def test_config4_full_size_properties(pkg, torch_mod):
    """n = 60 FCC spheres r = 0.12a, gamma = 3.46427 (just below the eps = 13 Pasteur limit
    sqrt(13)): fncB Hermitian and positive on random vectors; the shift-invert inner system
    (fncB - 0.3 fncA) y = x solved by the device MINRES to 1e-8 has that true residual."""
    say("config 4: n=60 sphere gamma=3.46427 properties")
    S = pkg.structures
    A = S.FCC_LATTICE
    dims = (60, 60, 60)
    theta = S.theta_of_k(S.fcc_band_path(8)[1], A)
    plan = pkg.Plan(dims, theta, S.bi_isotropic_table(13.0, 3.46427, A), S.sphere_ids(dims, 0.12, A))
    t = torch_mod
    x = t.as_tensor(crand(4 * plan.N), device="cuda")
    y = t.as_tensor(crand(4 * plan.N), device="cuda")
    bx, by = plan.fncB(x).reshape(-1), plan.fncB(y).reshape(-1)
    a, b = complex(t.vdot(y, bx)), complex(t.vdot(x, by))
    assert abs(a - np.conj(b)) <= 1e-12 * abs(a)
    assert complex(t.vdot(x, bx)).real > 0
    sol, iters, relres = plan.minres(x, sigma=0.3, tol=1e-8, maxit=20000)
    r = plan.fncB(sol) - 0.3 * plan.fncA(sol) - x
    true_rel = float(r.norm() / x.norm())
    say(f"  MINRES {int(iters[0])} iterations, estimate {relres[0]:.2e}, true {true_rel:.2e}")
    assert relres[0] <= 1e-8 and true_rel <= 1e-7
    plan.close()

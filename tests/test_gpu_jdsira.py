"""GPU parity of the JD-SIRA pieces: block projections, the device MINRES (the SIRA
inner solve) and the JD-SIRA driver, through the C-ABI, against the oracle.

Tolerances: projections 1e-12 relative; MINRES iterates against the oracle's
MINRES after the same number of steps 1e-10 relative (same algorithm, rounding
grows with the step count); eigenvalues 1e-10 relative to dense eigh; residual
||A u - w B u|| <= 10 * stop_tolerance (the driver's acceptance rule)."""
import numpy as np
import pytest
import scipy.linalg

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(77)


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


def woodpile_plan(pkg, fo, dims, thetas, gamma=0.2):
    S = pkg.structures
    table, ids = S.bi_isotropic_table(13.0, gamma, S.SC_LATTICE), S.woodpile_sc_ids(dims)
    plan = pkg.Plan(dims, thetas, table, ids)
    T = table[ids]
    probs = [fo.Problem(dims, th, T[:, 0], T[:, 1], T[:, 2]) for th in np.atleast_2d(thetas)]
    return plan, probs


THETAS = np.array([[0.4, -1.3, 2.2], [1.5, 0.1, -0.7]])


@pytest.mark.parametrize("m,length", [(1, 480), (7, 480), (12, 4 * 1000 + 4), (3, 4 * 13824)])
def test_block_dot_gemv(pkg, torch_mod, m, length):
    say(f"block_dot/gemv m={m} len={length}")
    t = torch_mod
    plan = pkg.Plan((4, 5, 6), THETAS[0], np.zeros((1, 3, 3, 3), complex) + np.eye(3), np.zeros(120, np.uint8))
    V = crand(m, length)
    w = crand(length)
    y = crand(m)
    base = crand(length)
    Vd, wd = t.as_tensor(V, device="cuda"), t.as_tensor(w, device="cuda")
    d = plan.block_dot(Vd, wd).cpu().numpy()
    assert rel(d, V.conj() @ w) < 1e-12
    g = plan.block_gemv(Vd, t.as_tensor(y)).cpu().numpy()
    assert rel(g, y @ V) < 1e-12
    out = t.as_tensor(base, device="cuda")
    plan.block_gemv(Vd, t.as_tensor(y), out=out, beta=-0.5)
    assert rel(out.cpu().numpy(), -0.5 * base + y @ V) < 1e-12
    plan.close()


@pytest.mark.parametrize("m,length", [(1, 480), (9, 4 * 1000 + 4), (40, 4 * 13824)])
def test_mgs_matches_sequential_and_orthogonalises(pkg, torch_mod, m, length):
    """fame_mgs = the column-by-column MGS of Locking4.m:326-330 (non-orthonormal V: same
    coefficients and result as the sequential loop), and against orthonormal columns the
    result is orthogonal to every column (a plausible index or conjugation slip fails one)."""
    say(f"mgs m={m} len={length}")
    t = torch_mod
    plan = pkg.Plan((4, 5, 6), THETAS[0], np.zeros((1, 3, 3, 3), complex) + np.eye(3), np.zeros(120, np.uint8))
    V = crand(m, length) / np.sqrt(length)
    x = crand(length)
    ref, cref = x.copy(), np.zeros(m, complex)
    for j in range(m):
        cref[j] = np.vdot(V[j], ref)
        ref = ref - cref[j] * V[j]
    td = t.as_tensor(x, device="cuda")
    c = plan.mgs(t.as_tensor(V, device="cuda"), td, coef=True)
    assert rel(td.cpu().numpy(), ref) < 1e-12
    assert rel(c.cpu().numpy(), cref) < 1e-12
    Q = np.linalg.qr(V.T)[0].T.copy()                       # orthonormal rows
    td = t.as_tensor(x, device="cuda")
    plan.mgs(t.as_tensor(Q, device="cuda"), td)
    r = td.cpu().numpy()
    assert np.abs(Q.conj() @ r).max() <= 1e-13 * np.linalg.norm(x)
    assert rel(r, x - Q.T @ (Q.conj() @ x)) < 1e-12
    plan.close()


def test_gram_schmidt_device_matches_mgs(pkg, torch_mod):
    """jdsira.gram_schmidt (block coefficients per step) = Gram_Schmidt.m's MGS loop,
    including the drop of a dependent vector (norm <= 1e-6 after orthogonalisation)."""
    say("gram_schmidt device vs MGS loop")
    t = torch_mod
    plan = pkg.Plan((4, 5, 6), THETAS[0], np.zeros((1, 3, 3, 3), complex) + np.eye(3), np.zeros(120, np.uint8))
    Y = crand(6, 4 * 1000)
    Y[3] = 0.5 * Y[0] - 2j * Y[1]                            # dependent: dropped
    ref = Y.copy()
    nq = 0
    for k in range(6):
        nr = np.linalg.norm(ref[k])
        if nr > 1e-6:
            ref[nq] = ref[k] / nr
            for j in range(k + 1, 6):
                ref[j] = ref[j] - np.vdot(ref[nq], ref[j]) * ref[nq]
            nq += 1
    Yd = t.as_tensor(Y, device="cuda")
    n_q = pkg.jdsira.gram_schmidt(plan, Yd)
    assert n_q == nq == 5
    got = Yd[:n_q].cpu().numpy()
    assert rel(got, ref[:nq]) < 1e-12
    assert np.abs(got.conj() @ got.T - np.eye(n_q)).max() < 1e-13
    plan.close()


@pytest.mark.parametrize("sigma", [0.0, 0.8])
def test_minres_iterates_match_oracle(pkg, fo, torch_mod, sigma):
    say(f"minres iterates sigma={sigma}")
    dims = (6, 5, 4)
    plan, probs = woodpile_plan(pkg, fo, dims, THETAS)
    x = crand(2, 4 * plan.N)
    for steps in (1, 4, 15):
        y, iters, relres = plan.minres(torch_mod.as_tensor(x), sigma=sigma, tol=1e-30, maxit=steps)
        y = y.cpu().numpy()
        for b in range(2):
            P = probs[b]
            yo, flag, ro, ito = fo.minres(lambda v: P.fncB(v) - sigma * P.fncA(v), x[b], 1e-30, steps)
            assert iters[b] == steps == ito
            assert rel(y[b], yo) < 1e-10, (steps, b)
            assert abs(relres[b] - ro) <= 1e-8 * ro
    plan.close()


@pytest.mark.parametrize("sigma", [0.0, 0.8])
def test_minres_stencil_partials_ragged_chunks(pkg, fo, torch_mod, sigma):
    """On a specialised grid MINRES takes u^H A u from the material pass's partial sums and
    u^H fncA u from its Lanczos update; with three k-points in chunks of two the last chunk's
    stencil leaves a different number of partials.  Iterates after 1, 4 and 15 steps equal the
    oracle's MINRES (1e-10), and the generic-stencil path (one more inner-product kernel) to
    1e-12."""
    say(f"minres stencil partials sigma={sigma}")
    dims = (8, 8, 8)
    thetas = np.vstack([THETAS, [[0.9, 0.6, -1.9]]])
    plan, probs = woodpile_plan(pkg, fo, dims, thetas)
    plan.set_chunk(2)
    x = crand(3, 4 * plan.N)
    for steps in (1, 4, 15):
        y, iters, relres = plan.minres(torch_mod.as_tensor(x), sigma=sigma, tol=1e-30, maxit=steps)
        y = y.cpu().numpy()
        plan.set_generic(True)
        yg, itg, _ = plan.minres(torch_mod.as_tensor(x), sigma=sigma, tol=1e-30, maxit=steps)
        plan.set_generic(False)
        for b in range(3):
            P = probs[b]
            yo, flag, ro, ito = fo.minres(lambda v: P.fncB(v) - sigma * P.fncA(v), x[b], 1e-30, steps)
            assert iters[b] == steps == ito == itg[b]
            assert rel(y[b], yo) < 1e-10, (steps, b)
            assert rel(y[b], yg[b].cpu().numpy()) < 1e-12, (steps, b)
            assert abs(relres[b] - ro) <= 1e-8 * ro
    plan.close()


def test_minres_converges(pkg, fo, torch_mod):
    say("minres converged solve, batched")
    dims = (8, 8, 8)
    plan, probs = woodpile_plan(pkg, fo, dims, THETAS)
    x = crand(2, 4 * plan.N)
    y, iters, relres = plan.minres(torch_mod.as_tensor(x), tol=1e-11, maxit=3000)
    y = y.cpu().numpy()
    for b in range(2):
        P = probs[b]
        assert relres[b] <= 1e-11 and 0 < iters[b] < 3000
        assert np.linalg.norm(P.fncB(y[b]) - x[b]) <= 1e-10 * np.linalg.norm(x[b])
    with pytest.raises(pkg.NotConverged):
        plan.minres(torch_mod.as_tensor(x), tol=1e-11, maxit=3, raise_on_maxit=True)
    y, iters, _ = plan.minres(torch_mod.as_tensor(x), tol=1e-11, maxit=3)
    assert list(iters) == [3, 3]
    plan.close()


@pytest.mark.parametrize("n", [8, 24])
def test_minres_finished_items_frozen(pkg, fo, torch_mod, n):
    """MINRES polls its state one batch of iterations behind, so iterations are queued after an
    item has converged; every pass then leaves the item alone.  A batch with a zero right-hand
    side (finished before the first iteration) and two items that converge at different steps:
    each item's iteration count and solution equal those of the item solved alone (within one
    iteration and 1e-9: the lone plan's stencil sums its partials in other groups), the zero
    item stays zero, and the solutions solve the shifted system to the tolerance."""
    say(f"minres finished items frozen n={n}")
    dims = (n, n, n)
    thetas = np.vstack([THETAS, [[0.9, 0.6, -1.9]]])
    plan, probs = woodpile_plan(pkg, fo, dims, thetas)
    x = crand(3, 4 * plan.N)
    x[0] = 0
    x[1] *= 1e-3   # the same tolerance relative to ||b|| converges at a different step per item
    sigma, tol = 0.3, 1e-10
    y, iters, relres = plan.minres(torch_mod.as_tensor(x), sigma=sigma, tol=tol, maxit=4000)
    y = y.cpu().numpy()
    assert iters[0] == 0 and not y[0].any()
    for b in (1, 2):
        p1, _ = woodpile_plan(pkg, fo, dims, thetas[b])
        y1, it1, _ = p1.minres(torch_mod.as_tensor(x[b]), sigma=sigma, tol=tol, maxit=4000)
        assert abs(int(it1[0]) - int(iters[b])) <= 1, (b, it1, iters)
        assert rel(y[b], y1.cpu().numpy()) < 1e-9, b
        p1.close()
        if n == 8:
            P = probs[b]
            r = P.fncB(y[b]) - sigma * P.fncA(y[b]) - x[b]
            assert np.linalg.norm(r) <= 10 * tol * np.linalg.norm(x[b])
        assert relres[b] <= tol
    plan.close()


def test_jdsira_woodpile_matches_dense(pkg, fo, torch_mod):
    say("JD-SIRA woodpile n=6 vs dense GEP")
    dims = (6, 6, 6)
    S = pkg.structures
    theta = S.theta_of_k(np.pi * np.array([0.5, 0.3, 0.1]), S.SC_LATTICE)
    plan, (P,) = woodpile_plan(pkg, fo, dims, theta)
    n4 = 4 * P.N
    Ad = np.stack([P.fncA(e) for e in np.eye(n4, dtype=complex)], 1)
    Bd = np.stack([P.fncB(e) for e in np.eye(n4, dtype=complex)], 1)
    mu = scipy.linalg.eigh(Ad, Bd, eigvals_only=True)
    k = 4
    want = np.sort(1.0 / mu[mu > 0])[:k]
    res = pkg.jdsira.GEP_AB_Herm_JDSIRA_Driver4(plan, k, stop_tolerance=1e-8, seed=0)
    got = np.sort(res.ew.real)
    say(f"  omega {got} inner {res.total_inner} outer {res.outer_iterations}")
    assert np.allclose(got, want, rtol=1e-10)
    EV = res.ev.cpu().numpy()
    for j in range(k):
        u = EV[j]
        assert abs(np.linalg.norm(u) - 1) < 1e-12
        assert np.linalg.norm(Bd @ u - res.ew[j] * (Ad @ u)) <= 1e-7
    plan.close()


def test_fused_jd_steps_match_composed_primitives(pkg, fo, torch_mod):
    """fame_jd_ritz_residual and fame_jd_expand (one outer JD-SIRA step in two calls) give what
    the separate device primitives give: u = V y / ||V y||, r = fncB u - theta fncA u, ||r||;
    MINRES, the tiny-norm rescale, MGS, normalisation and the two projection columns."""
    say("fused JD-SIRA steps == composed primitives, n=8")
    t = torch_mod
    dims = (8, 8, 8)
    plan, _ = woodpile_plan(pkg, fo, dims, THETAS[0])
    L = 4 * plan.N
    m = 5
    V = t.zeros((m + 2, L), dtype=t.complex128, device="cuda")
    V[:m] = t.as_tensor(crand(m, L), device="cuda")
    for j in range(m):  # orthonormal columns, as JD-SIRA keeps them
        plan.mgs(V[:j], V[j]) if j else None
        V[j] /= V[j].norm()
    y = crand(m)
    theta = 1.37 + 0.0j
    u, r, rn = plan.jd_ritz_residual(V[:m], y, theta)
    u0 = plan.block_gemv(V[:m], t.as_tensor(y))
    u0 = u0 / u0.norm()
    r0 = plan.fncB(u0) - theta * plan.fncA(u0)
    assert float((u - u0).norm()) < 1e-13 and float((r - r0).norm() / r0.norm()) < 1e-13
    assert abs(rn - float(r0.norm())) < 1e-12 * rn
    its, mcol, lcol = plan.jd_expand(V, m, r, 0.3, 1e-2, 500)
    tv, its0, _ = plan.minres(r0, sigma=0.3, tol=1e-2, maxit=500)
    tv = tv.contiguous()
    plan.mgs(V[:m], tv)
    tv = tv / tv.norm()
    assert its == int(its0[0]) and float((V[m] - tv).norm()) < 1e-11
    assert np.abs(mcol - plan.block_dot(V[:m + 1], plan.fncB(V[m])).cpu().numpy()).max() < 1e-12 * np.abs(mcol).max()
    assert np.abs(lcol - plan.block_dot(V[:m + 1], plan.fncA(V[m])).cpu().numpy()).max() < 1e-12 * np.abs(lcol).max()
    plan.close()

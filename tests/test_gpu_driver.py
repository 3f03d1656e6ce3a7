"""GPU: device residual evaluation (EvaluteError_Chiral_no_inv) against the oracle with the
explicit sparse curl, and the band-structure driver (plain and batched) end to end."""
import numpy as np
import pytest
import scipy.io

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(91)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_residuals_match_oracle(pkg, fo, torch_mod):
    print("EvaluteError_Chiral_no_inv vs oracle", flush=True)
    dims = (6, 5, 4)
    S = pkg.structures
    thetas = np.array([[0.4, -1.3, 2.2], [1.5, 0.1, -0.7]])
    table, ids = S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE), S.woodpile_sc_ids(dims)
    plan = pkg.Plan(dims, thetas, table, ids)
    N = plan.N
    ev = crand(2, 3, 4 * N)          # not eigenvectors: O(1) residuals, compared to 1e-12
    omega = rng.uniform(0.5, 2, (2, 3))
    e_h = crand(2, 3, 6 * N)
    rn, ro = plan.EvaluteError_Chiral_no_inv(torch_mod.as_tensor(ev), omega, torch_mod.as_tensor(e_h))
    T = table[ids]
    for b in range(2):
        P = fo.Problem(dims, thetas[b], T[:, 0], T[:, 1], T[:, 2])
        C = fo.curl_C(dims, P.delta, thetas[b])
        e = e_h[b][:, :3 * N].T
        h = e_h[b][:, 3 * N:].T
        r_nf, r_org = fo.evaluate_error(ev[b].T, omega[b], e, h, C, P.fncA, P.fncB, P.mtx_B)
        assert np.allclose(rn[b], r_nf, rtol=1e-12, atol=0)
        assert np.allclose(ro[b], r_org, rtol=1e-12, atol=0)
    plan.close()


def test_band_structure_driver_plain_and_batched(pkg, torch_mod, tmp_path):
    print("band-structure driver n=8, 3 k", flush=True)
    dims = (8, 8, 8)
    S = pkg.structures
    ks = S.sc_band_path(part_num=2)[:3]
    thetas = np.stack([S.theta_of_k(k, S.SC_LATTICE) for k in ks])
    table, ids = S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE), S.woodpile_sc_ids(dims)
    path = tmp_path / "band.mat"
    rb = pkg.driver.band_structure_lanczos(pkg, dims, thetas, table, ids, nev=4, batched=True, mat_path=str(path))
    rp = pkg.driver.band_structure_lanczos(pkg, dims, thetas, table, ids, nev=4, batched=False)
    assert np.allclose(rb["Freq_mtx"], rp["Freq_mtx"], rtol=1e-10)
    assert rb["res_NFSEP"].max() < 1e-8 and rp["res_NFSEP"].max() < 1e-8
    assert rb["res_ORGSEP"].max() < 1e-6
    m = scipy.io.loadmat(str(path))
    assert np.allclose(m["Freq_mtx"], rb["Freq_mtx"])
    assert rb["not_converged"] == [] and rp["not_converged"] == []


def test_lanczos_reports_unconverged_items(pkg, torch_mod):
    """maxiter = 1 restart with a tight tolerance leaves the k-points unconverged: the call raises
    NotConverged by default; with raise_on_maxit=False it returns every item's last Ritz pairs
    (finite, ev written) with iters negated, and a full solve afterwards reports positive steps."""
    print("Lanczos per-item convergence flags, n=6, 2 k", flush=True)
    dims = (6, 6, 6)
    S = pkg.structures
    thetas = np.stack([S.theta_of_k(np.pi * np.array(k), S.SC_LATTICE) for k in ([0.5, 0.3, 0.1], [0.2, 0.4, 0.6])])
    plan = pkg.Plan(dims, thetas, S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE), S.woodpile_sc_ids(dims))
    with pytest.raises(pkg.NotConverged):
        plan.Lanczos(4, tol=1e-14, maxiter=1)
    ev, ew, it = plan.Lanczos(4, tol=1e-14, maxiter=1, raise_on_maxit=False)
    assert np.all(it < 0) and np.all(np.isfinite(ew)) and np.all(ew != 0)
    assert bool(torch_mod.isfinite(ev).all()) and float(ev.abs().max()) > 0
    ev, ew, it = plan.Lanczos(4, tol=1e-10, maxiter=10000, raise_on_maxit=False)
    assert np.all(it > 0)
    plan.close()


def test_lanczos_refuses_bad_inner_solver_arguments(pkg, torch_mod):
    """cg_maxit < 1 or cg_tol <= 0 is an argument error, raised before any inner solve."""
    print("Lanczos refuses cg_maxit = 0 and cg_tol = 0", flush=True)
    dims = (6, 6, 6)
    S = pkg.structures
    theta = S.theta_of_k(np.pi * np.array([0.5, 0.3, 0.1]), S.SC_LATTICE)
    plan = pkg.Plan(dims, theta, S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE), S.woodpile_sc_ids(dims))
    for kw in ({"cg_maxit": 0}, {"cg_tol": 0.0}):
        with pytest.raises(pkg.FameError):
            plan.Lanczos(2, tol=1e-8, maxiter=10, **kw)
    ev, ew, it = plan.Lanczos(2, tol=1e-8, maxiter=1000)  # the plan still works afterwards
    assert np.all(it > 0)
    plan.close()


def test_lanczos_wide_restart_matches_narrow(pkg, torch_mod):
    """A restart basis wider than 32 columns (p = 34) forms Q U through a scratch product instead
    of in place; it converges to the same eigenvalues as p = 20 (1e-10), with small residuals."""
    print("Lanczos p = 34 (scratch restart product) vs p = 20 (in place), n=6", flush=True)
    dims = (6, 6, 6)
    S = pkg.structures
    theta = S.theta_of_k(np.pi * np.array([0.5, 0.3, 0.1]), S.SC_LATTICE)
    plan = pkg.Plan(dims, theta, S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE), S.woodpile_sc_ids(dims))
    _, ew20, _ = plan.Lanczos(12, tol=1e-12, maxiter=10000, p=20)
    ev34, ew34, it34 = plan.Lanczos(12, tol=1e-12, maxiter=10000, p=34)
    assert it34[0] > 34  # at least one restart
    assert np.allclose(np.sort(ew34[0]), np.sort(ew20[0]), rtol=1e-10, atol=0)
    for j in range(12):
        u = ev34[0, j].contiguous()
        r = plan.fncB(u) - float(1.0 / ew34[0, j]) * plan.fncA(u)
        assert float(r.norm() / u.norm()) < 1e-8
    plan.close()


def test_batched_lanczos_two_streams_repeatable(pkg, torch_mod):
    """A batched Lanczos solve with its fncB and fused-CG chunks dealt over two streams gives the
    same Lanczos step counts, eigenvalues and eigenvectors, bit for bit, when it is repeated on
    the same plan (k-points finish at different steps, so the solve also runs on packed views)."""
    print("batched Lanczos, two streams, repeated solve bit for bit, n=24, 12 k", flush=True)
    dims = (24, 24, 24)
    S = pkg.structures
    ks = S.sc_band_path(3)[:12]
    thetas = np.stack([S.theta_of_k(k, S.SC_LATTICE) for k in ks])
    plan = pkg.Plan(dims, thetas, S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE), S.woodpile_sc_ids(dims))
    plan.set_chunk(2)
    plan.set_streams(2)
    runs = []
    for _ in range(3):
        ev, ew, it = plan.Lanczos(6, tol=1e-10, maxiter=10000, p=16, cg_tol=1e-12, cg_maxit=10000)
        runs.append((ev.cpu().numpy(), ew.copy(), np.abs(it).copy()))
    assert len(set(runs[0][2].tolist())) > 1  # items finish at different steps
    for ev, ew, it in runs[1:]:
        assert np.array_equal(it, runs[0][2])
        assert np.array_equal(ew, runs[0][1])
        assert np.array_equal(ev, runs[0][0])
    plan.close()


def test_distinct_plans_from_threads_match_sequential(pkg, torch_mod):
    """Plans are independent handles: two plans (different media) driven at once from two host
    threads, each on its own stream, give the results they give one after another, bit for bit
    (Lanczos with its fused CG, and a JD-SIRA solve)."""
    import threading

    print("two plans from two threads == sequential, n=8", flush=True)
    dims = (8, 8, 8)
    S = pkg.structures
    theta = S.theta_of_k(np.pi * np.array([0.5, 0.3, 0.1]), S.SC_LATTICE)
    ids = S.woodpile_sc_ids(dims)
    tables = [S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE), S.bi_isotropic_table(9.0, 0.5, S.SC_LATTICE)]

    def solve(i, out, stream):
        with torch_mod.cuda.stream(stream):
            pl = pkg.Plan(dims, theta, tables[i], ids)
            ev, ew, it = pl.Lanczos(4, tol=1e-10, maxiter=10000, p=8)
            r = pkg.jdsira.GEP_AB_Herm_JDSIRA_Driver4(pl, 3, stop_tolerance=1e-9, target=2.0, target_type="CL",
                                                       sigma_inner=2.0, seed=0)
            stream.synchronize()
            out[i] = (ew.copy(), it.copy(), np.sort(r.ew.real), r.outer_iterations)
            pl.close()

    streams = [torch_mod.cuda.Stream() for _ in tables]
    seq = [None, None]
    for i in range(2):
        solve(i, seq, streams[i])
    par = [None, None]
    th = [threading.Thread(target=solve, args=(i, par, streams[i])) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i in range(2):
        assert par[i] is not None
        assert np.array_equal(par[i][0], seq[i][0]) and np.array_equal(par[i][1], seq[i][1])
        assert np.array_equal(par[i][2], seq[i][2]) and par[i][3] == seq[i][3]

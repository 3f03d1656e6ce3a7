"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the oracles):
the oracles must keep reproducing them (CPU), and the HIP path must match them (GPU).
woodpile_*: SC operators, material apply and Lanczos; fcc_*: FCC T, T^H, NFGEP matvec and
eigenvector map; vertex_*: per-vertex media on a skewed lattice."""
import glob
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "woodpile_*.npz")))
FCC_FILES = sorted(glob.glob(os.path.join(HERE, "golden", "fcc_*.npz")))
VERTEX_FILES = sorted(glob.glob(os.path.join(HERE, "golden", "vertex_*.npz")))


def _make_golden():
    """The fixture generator as a module (for the closed-form NFGEP coefficients)."""
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def load(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_golden(fo, path):
    g = load(path)
    dims = tuple(int(d) for d in g["dims"])
    T = g["table"][g["ids"]]
    P = fo.Problem(dims, g["theta"], T[:, 0], T[:, 1], T[:, 2])
    f = P.fnc
    assert rel(P.fncB(g["x4"]), g["fncB"]) < 1e-13
    assert rel(P.fncA(g["x4"]), g["fncA"]) < 1e-15
    assert rel(f.Psi(g["x3"]), g["Psi"]) < 1e-13
    assert rel(f.Prs(g["x3"]), g["Prs"]) < 1e-13
    assert rel(f.Qr(g["x2"]), g["Qr"]) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_gpu_matches_golden(pkg, path):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print(f"golden {os.path.basename(path)}", flush=True)
    g = load(path)
    dims = tuple(int(d) for d in g["dims"])
    plan = pkg.Plan(dims, g["theta"], g["table"], g["ids"])
    T = torch.as_tensor
    c = lambda t: t.cpu().numpy()  # noqa: E731
    assert rel(c(plan.fncB(T(g["x4"]))), g["fncB"]) < 1e-12
    assert rel(c(plan.fncA(T(g["x4"]))), g["fncA"]) < 1e-12
    assert rel(c(plan.MtxVecMul_FFT(T(g["x3"]), "notranspose")), g["Psi"]) < 1e-12
    assert rel(c(plan.MtxVecMul_FFT(T(g["x3"]), "transpose")), g["Psis"]) < 1e-12
    assert rel(c(plan.Pr(T(g["x2"]))), g["Pr"]) < 1e-12
    assert rel(c(plan.Qr(T(g["x2"]))), g["Qr"]) < 1e-12
    assert rel(c(plan.Prs(T(g["x3"]))), g["Prs"]) < 1e-12
    assert rel(c(plan.Qrs(T(g["x3"]))), g["Qrs"]) < 1e-12
    assert rel(c(plan.material_apply(T(g["x6"]))), g["material"]) < 1e-12
    ev, ew, _ = plan.Lanczos(4, tol=1e-10, maxiter=200)
    assert np.allclose(np.sort(ew[0])[::-1], g["lanczos_ew"], rtol=1e-10)
    plan.close()


@pytest.mark.parametrize("path", FCC_FILES, ids=[os.path.basename(f) for f in FCC_FILES])
def test_fcc_oracle_reproduces_golden(path):
    from oracle import fcc_oracle as fc

    g = load(path)
    n1, n2, n3 = (int(d) for d in g["dims"])
    tab = fc.fft_tables(n1, n2, n3, g["vec_k"], *g["lattice"], g["delta"])
    svd = fc.svd_curl(tab)
    xi, zeta, phi = _make_golden().fcc_coefficients(n1 * n2 * n3)
    assert rel(fc.T(g["u"], tab), g["T"]) < 1e-13
    assert rel(fc.TH(g["u"], tab), g["TH"]) < 1e-13
    assert rel(fc.nfgep_A_matvec(g["x32"], tab, svd, xi, zeta, phi), g["nfgep"]) < 1e-13
    if "eigenvector" in g:
        assert rel(fc.compute_eigenvector(g["x32"], tab, svd, xi, zeta, phi), g["eigenvector"]) < 1e-13


@pytest.mark.parametrize("path", VERTEX_FILES, ids=[os.path.basename(f) for f in VERTEX_FILES])
def test_vertex_oracle_reproduces_golden(fo, path):
    g = load(path)
    dims = tuple(int(d) for d in g["dims"])
    Nv, Mv, Jv = fo.vertex_material(g["eps"], g["gamma"], float(g["mu"]), g["lattice"])
    P = fo.Problem(dims, g["theta"], Nv, Mv, Jv)
    assert rel(P.fncB(g["x4"]), g["fncB"]) < 1e-13
    assert rel(P.fncA(g["x4"]), g["fncA"]) < 1e-15


@pytest.mark.gpu
@pytest.mark.parametrize("path", FCC_FILES, ids=[os.path.basename(f) for f in FCC_FILES])
def test_gpu_fcc_matches_golden(pkg, path):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print(f"golden {os.path.basename(path)}", flush=True)
    g = load(path)
    dims = tuple(int(d) for d in g["dims"])
    n = int(np.prod(dims))
    plan = pkg.FccPlan(dims, g["vec_k"], tuple(g["lattice"]), g["delta"])
    T = torch.as_tensor
    c = lambda t: t.cpu().numpy()  # noqa: E731
    xi, zeta, phi = (T(v) for v in _make_golden().fcc_coefficients(n))
    assert rel(c(plan.T(T(g["u"][None, :]))[0]), g["T"]) < 1e-12
    assert rel(c(plan.TH(T(g["u"][None, :]))[0]), g["TH"]) < 1e-12
    assert rel(c(plan.mtx_prod_vec_A_NFGEP(T(g["x32"]), xi, zeta, phi)), g["nfgep"]) < 1e-12
    if "eigenvector" in g:
        assert rel(c(plan.Compute_eigenvector(T(g["x32"]), xi, zeta, phi)), g["eigenvector"]) < 1e-12
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("path", VERTEX_FILES, ids=[os.path.basename(f) for f in VERTEX_FILES])
def test_gpu_vertex_matches_golden(pkg, path):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print(f"golden {os.path.basename(path)}", flush=True)
    g = load(path)
    dims = tuple(int(d) for d in g["dims"])
    plan = pkg.Plan(dims, g["theta"], g["table"], g["ids"])
    plan.set_vertex_material(g["eps"], g["gamma"], float(g["mu"]), g["lattice"])
    T = torch.as_tensor
    assert rel(plan.fncB(T(g["x4"])).cpu().numpy(), g["fncB"]) < 1e-12
    assert rel(plan.fncA(T(g["x4"])).cpu().numpy(), g["fncA"]) < 1e-12
    plan.close()

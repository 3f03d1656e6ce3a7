"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the oracle):
the oracle must keep reproducing them (CPU), and the HIP path must match them (GPU)."""
import glob
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


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

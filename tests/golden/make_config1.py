"""Write tests/golden/config1_n24_omega.json: the ten smallest positive omega of the
reference's CPU-runnable case (SC woodpile n = 24, eps = 13, mu = 1, gamma = 0.2, one k-point
pi (0.5, 0.3, 0.1)) from the oracle's restatement of Lanczos.m (p = 20, tol 1e-10, pcg tol
1e-12, v0 = ones as band_structure_woodpile_Lanczos_n60_gamma0p2.m:78).

The reference (MATLAB) cannot run here, so the fixture is checked, before it is written,
against a second algorithm on the same operators: ARPACK (scipy eigsh, generalized mode 2 on
fncA x = mu fncB x with fncB^{-1} by pcg), which must agree to 1e-10.  Run from the repo root
in the build container (a few minutes on the host cores):
    python tests/golden/make_config1.py
"""
import json
import os
import sys
import time

import numpy as np
import scipy.sparse.linalg as spl

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import fame_oracle as fo  # noqa: E402


def main():
    import __graft_entry__ as ge

    S = ge.load_package().structures
    n, nev, p = 24, 10, 20
    dims = (n, n, n)
    kk = np.array([0.5, 0.3, 0.1])
    theta = S.theta_of_k(np.pi * kk, S.SC_LATTICE)
    table = S.bi_isotropic_table(13.0, 0.2, S.SC_LATTICE)
    ids = S.woodpile_sc_ids(dims)
    T = table[ids]
    P = fo.Problem(dims, theta, T[:, 0], T[:, 1], T[:, 2])
    L = 4 * P.N
    t0 = time.time()
    binv = lambda x: P.fncinvB(x, 1e-12, 10000)[0]  # noqa: E731
    _, ew, steps = fo.lanczos(P.fncA, P.fncB, binv, L, nev, 1e-10, 10000, p, np.ones(L, complex))
    omega = np.sort(1.0 / np.real(ew))
    print(f"oracle Lanczos: {steps} steps, {time.time() - t0:.0f} s, omega {omega}", flush=True)
    t0 = time.time()
    A = spl.LinearOperator((L, L), matvec=P.fncA, dtype=complex)
    B = spl.LinearOperator((L, L), matvec=P.fncB, dtype=complex)
    Bi = spl.LinearOperator((L, L), matvec=binv, dtype=complex)
    mu = spl.eigsh(A, k=nev, M=B, Minv=Bi, which="LA", tol=1e-13, return_eigenvectors=False)
    om_arpack = np.sort(1.0 / mu)
    diff = float(np.max(np.abs(om_arpack - omega) / omega))
    print(f"ARPACK: {time.time() - t0:.0f} s, omega {om_arpack}, max rel diff {diff:.2e}", flush=True)
    if not diff < 1e-10:
        raise SystemExit("oracle Lanczos and ARPACK disagree; fixture not written")
    out = {"case": "SC woodpile n=24, eps=13, mu=1, gamma=0.2, k = pi (0.5, 0.3, 0.1), theta = k . a_d",
           "solver": "oracle Lanczos (fun_Lanczos/Lanczos.m restated), p=20, tol 1e-10, pcg tol 1e-12, v0 = ones",
           "lanczos_steps": int(steps), "omega": omega.tolist(), "omega_arpack": om_arpack.tolist(),
           "max_rel_diff_lanczos_vs_arpack": diff}
    with open(os.path.join(HERE, "config1_n24_omega.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()

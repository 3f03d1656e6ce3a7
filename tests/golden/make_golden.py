"""Write the small golden fixtures under tests/golden/ from the CPU oracle.

The reference (MATLAB) ships no vectors and cannot run here, so these fixtures
are the oracle's outputs on seeded inputs; they freeze the oracle (regression)
and give the GPU tests fixed expected outputs.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import fame_oracle as fo  # noqa: E402


def woodpile_table_ids(dims, gamma):
    import __graft_entry__ as ge

    S = ge.load_package().structures
    return S.bi_isotropic_table(13.0, gamma, S.SC_LATTICE), S.woodpile_sc_ids(dims)


CASES = {
    # name: (dims, k/pi, gamma)
    "woodpile_n6_k053": ((6, 6, 6), (0.5, 0.3, 0.1), 0.2),
    "woodpile_4x5x6_k1": ((4, 5, 6), (1.0, -0.4, 0.25), 0.8),
}


def make(name, dims, kk, gamma, rng):
    table, ids = woodpile_table_ids(dims, gamma)
    theta = np.pi * np.asarray(kk, float)  # SC lattice: theta = k
    T = table[ids]
    P = fo.Problem(dims, theta, T[:, 0], T[:, 1], T[:, 2])
    N = P.N
    cr = lambda n: rng.standard_normal(n) + 1j * rng.standard_normal(n)  # noqa: E731
    x4, x3, x2, x6 = cr(4 * N), cr(3 * N), cr(2 * N), cr(6 * N)
    f = P.fnc
    B = P.mtx_B
    h, e = x6[:3 * N], x6[3 * N:]
    mat = np.concatenate([B["M_int"] @ h + B["J_int"].conj().T @ e, B["J_int"] @ h + B["N_int"] @ e])
    k = 4
    ev, ew, _ = fo.lanczos(P.fncA, P.fncB, lambda r: P.fncinvB(r)[0], 4 * N, k, 1e-10, 200, 2 * k,
                           np.ones(4 * N, complex))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), dims=np.array(dims), theta=theta, table=table, ids=ids,
                        x4=x4, x3=x3, x2=x2, x6=x6, fncB=P.fncB(x4), fncA=P.fncA(x4),
                        Psi=f.Psi(x3), Psis=f.Psis(x3), Pr=f.Pr(x2), Qr=f.Qr(x2), Prs=f.Prs(x3), Qrs=f.Qrs(x3),
                        material=mat, lanczos_ew=np.sort(ew.real)[::-1])


if __name__ == "__main__":
    rng = np.random.default_rng(20261203)
    for name, (dims, kk, g) in CASES.items():
        make(name, dims, kk, g, rng)
        print("wrote", name)

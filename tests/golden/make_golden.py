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


FCC_CASES = {
    # name: (dims, vec_k); (8, 8, 8) is a specialised cubic length (the fused NFGEP kernels)
    "fcc_4x6x6": ((4, 6, 6), (0.31, -0.17, 0.23)),
    "fcc_n8": ((8, 8, 8), (0.12, 0.44, -0.3)),
}


def fcc_coefficients(n):
    """Diagonal xi, zeta, Phi (24n each) of the NFGEP fixtures: closed-form in the index, so
    the fixture files need not hold them (tests/test_golden.py computes the same)."""
    i = np.arange(24 * n, dtype=float)
    xi = 0.4 + 0.3j + 0.2 * np.cos(0.37 * i) + 0.1j * np.sin(0.11 * i)
    zeta = np.conj(xi) + 0.05 * np.sin(0.23 * i)
    phi = 3.0 + np.cos(0.05 * i) + 0.5j * np.sin(0.07 * i)
    return xi, zeta, phi


def make_fcc(name, dims, vec_k, rng, eigenvector=True):
    """FCC skewed-grid operators: T and T^H of one vector, the NFGEP matvec and (optionally)
    the eigenvector map with diagonal xi, zeta, Phi (fcc_oracle)."""
    from oracle import fcc_oracle as fc

    n1, n2, n3 = dims
    lat = np.stack(fc.fcc_lattice())
    delta = np.array([1 / n1, np.sqrt(3) / 2 / n2, np.sqrt(6) / 3 / n3])
    tab = fc.fft_tables(n1, n2, n3, np.asarray(vec_k, float), *lat, delta)
    svd = fc.svd_curl(tab)
    n = n1 * n2 * n3
    cr = lambda m: rng.standard_normal(m) + 1j * rng.standard_normal(m)  # noqa: E731
    u = cr(n)
    x32 = cr(32 * n)
    xi, zeta, phi = fcc_coefficients(n)
    extra = {"eigenvector": fc.compute_eigenvector(x32, tab, svd, xi, zeta, phi)} if eigenvector else {}
    np.savez_compressed(os.path.join(HERE, name + ".npz"), dims=np.array(dims), vec_k=np.asarray(vec_k, float),
                        lattice=lat, delta=delta, u=u, T=fc.T(u, tab), TH=fc.TH(u, tab), x32=x32,
                        nfgep=fc.nfgep_A_matvec(x32, tab, svd, xi, zeta, phi), **extra)


def make_vertex(name, dims, kk, rng):
    """Per-vertex bi-isotropic media (eps_v ~ U[1, 13] in the woodpile rods, gamma_v = 0.3,
    mu = 1.4, a skewed lattice): fncB and fncA."""
    table, ids = woodpile_table_ids(dims, 0.2)
    theta = np.pi * np.asarray(kk, float)
    A = np.array([[1.0, 0.5, 0.2], [0.0, 0.9, 0.3], [0.1, 0.0, 1.1]])
    eps = np.where(ids == 1, rng.uniform(1, 13, ids.size), 1.0)
    gam = np.where(ids == 1, 0.3, 0.0)
    mu = 1.4
    Nv, Mv, Jv = fo.vertex_material(eps, gam, mu, A)
    P = fo.Problem(dims, theta, Nv, Mv, Jv)
    x4 = rng.standard_normal(4 * P.N) + 1j * rng.standard_normal(4 * P.N)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), dims=np.array(dims), theta=theta, table=table, ids=ids,
                        eps=eps, gamma=gam, mu=mu, lattice=A, x4=x4, fncB=P.fncB(x4), fncA=P.fncA(x4))


if __name__ == "__main__":
    rng = np.random.default_rng(20261203)
    for name, (dims, kk, g) in CASES.items():
        make(name, dims, kk, g, rng)
        print("wrote", name)
    rng = np.random.default_rng(20261204)
    for name, (dims, vk) in FCC_CASES.items():
        make_fcc(name, dims, vk, rng, eigenvector=(name == "fcc_4x6x6"))
        print("wrote", name)
    make_vertex("vertex_n6_skew", (6, 6, 6), (0.5, 0.3, 0.1), np.random.default_rng(20261205))
    print("wrote vertex_n6_skew")

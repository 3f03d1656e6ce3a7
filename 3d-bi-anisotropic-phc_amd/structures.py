"""Per-structure setup on the host: material tensors, vertex voxelisation, k-paths.

This is the operand producer in front of the device path (the reference's
``Physical_settings_*``, ``Material_Matrix_Generate_Mult_Point`` and the
driver's k loop).  It yields the inputs ``Plan`` takes: ``theta`` (nb,3), a
material table (n_mat,3,3,3) and per-vertex material ids.
"""
from __future__ import annotations

import numpy as np

SC_LATTICE = np.eye(3)
# FCC primitive vectors as columns (InputData/BS_Gyroid/Physical_settings_woodpile.m:7)
FCC_LATTICE = 0.5 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])


def material_slots(eps, mu, xi, zeta, lattice_vec):
    """N-, M-, J-slot tensors of one medium, shape (3,3,3) complex.

    Inverts the 6x6 [eps xi; zeta mu] and keeps its blocks (Physical_settings_woodpile.m:91-96),
    then the covariant form A' X A (…:184-194); slot map N <- (1:3,1:3), M <- (4:6,4:6),
    J <- (1:3,4:6) (Public/Material_Matrix_Generate_Mult_Point.m:28-31).
    """
    big = np.block([[np.asarray(eps, complex), np.asarray(xi, complex)],
                    [np.asarray(zeta, complex), np.asarray(mu, complex)]])
    inv = np.linalg.inv(big)
    A = np.asarray(lattice_vec, float)
    cov = lambda X: A.T @ (X @ A)  # noqa: E731
    return np.stack([cov(inv[:3, :3]), cov(inv[3:, 3:]), cov(inv[:3, 3:])])


def bi_isotropic_table(eps_in, gamma, lattice_vec, mu_in=1.0):
    """Table [vacuum, inclusion] for eps=eps_in*I, mu=mu_in*I, xi=i*gamma*I, zeta=-i*gamma*I."""
    I3 = np.eye(3)
    vac = material_slots(I3, I3, 0 * I3, 0 * I3, lattice_vec)
    inc = material_slots(eps_in * I3, mu_in * I3, 1j * gamma * I3, -1j * gamma * I3, lattice_vec)
    return np.stack([vac, inc])


def vertex_coords(dims):
    """Lattice coordinates (i1/n1, i2/n2, i3/n3) of every vertex, x fastest (Grid_Generate_Mult_Point.m:7-17)."""
    n1, n2, n3 = dims
    i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    return np.stack([i1.ravel() / n1, i2.ravel() / n2, i3.ravel() / n3], axis=1)


def woodpile_sc_ids(dims, width=0.25, a=1.0):
    """Woodpile in a simple-cubic cell: 4 rod layers of height a/4 along z.

    Layer 1 (z in [0,1/4)): rod along x centred at y = 1/4; layer 2: rod along y
    centred at x = 1/4; layer 3: along x centred at y = 3/4 (layer 1 shifted 1/2
    in y); layer 4: along y centred at x = 3/4.  Rod cross-section width x height
    = width x a/4.  Distances along the rod-normal axis are periodic.
    Returns uint8 ids (1 inside, 0 outside).
    """
    r = vertex_coords(dims) * a
    x, y, z = r[:, 0], r[:, 1], r[:, 2]
    hw = 0.5 * width

    def near(u, c):
        d = np.abs(((u - c + 0.5 * a) % a) - 0.5 * a)
        return d < hw

    layer = np.floor(z / (0.25 * a)).astype(int)
    inside = np.zeros(x.shape, bool)
    inside |= (layer == 0) & near(y, 0.25 * a)
    inside |= (layer == 1) & near(x, 0.25 * a)
    inside |= (layer == 2) & near(y, 0.75 * a)
    inside |= (layer == 3) & near(x, 0.75 * a)
    return inside.astype(np.uint8)


def sphere_ids(dims, radius, lattice_vec, center_coef=(0.5, 0.5, 0.5)):
    """Sphere of given radius at lattice coordinates ``center_coef`` (Physical_settings_sphere sphere_isofun).

    Point r = coef * A' is inside when |r - c| < radius; periodic images are
    included (minimum image over the 27 neighbouring cells).
    """
    A = np.asarray(lattice_vec, float)
    coef = vertex_coords(dims)
    c = np.asarray(center_coef, float)
    best = np.full(coef.shape[0], np.inf)
    for s in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).reshape(3, -1).T:
        d = (coef - c + s) @ A.T
        best = np.minimum(best, np.linalg.norm(d, axis=1))
    return (best < radius).astype(np.uint8)


def tetra_ids(dims, edge, lattice_vec, center_coef=(0.5, 0.5, 0.5)):
    """Regular tetrahedron of given edge centred at lattice coordinates ``center_coef``.

    Vertices at c + edge/sqrt(8) * (+-1,+-1,+-1) with an even number of minus
    signs; inside = all four face half-space tests (closed), minimum image.
    """
    A = np.asarray(lattice_vec, float)
    coef = vertex_coords(dims)
    c = np.asarray(center_coef, float)
    s = edge / np.sqrt(8.0)
    V = s * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    faces = []
    for i in range(4):
        o = [V[j] for j in range(4) if j != i]
        nrm = np.cross(o[1] - o[0], o[2] - o[0])
        if nrm @ (V[i] - o[0]) < 0:
            nrm = -nrm
        faces.append((nrm, o[0]))
    inside = np.zeros(coef.shape[0], bool)
    for sh in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).reshape(3, -1).T:
        d = (coef - c + sh) @ A.T
        ok = np.ones(coef.shape[0], bool)
        for nrm, o in faces:
            ok &= (d - o) @ nrm >= -1e-12
        inside |= ok
    return inside.astype(np.uint8)


def theta_of_k(k, lattice_vec):
    """theta_d = k . a_d, with k = 0 nudged by +1e-4 (band_structure_woodpile_Lanczos_n60_gamma0p2.m:49-56)."""
    k = np.asarray(k, float).copy()
    if np.linalg.norm(k) == 0:
        k = k + 1e-4
    A = np.asarray(lattice_vec, float)
    return np.array([k @ A[:, 0], k @ A[:, 1], k @ A[:, 2]])


def sc_band_path(part_num=10):
    """Gamma-X-M-R-Gamma of the simple-cubic zone as the reference builds paths
    (linspace per segment with shared endpoints once, Physical_settings_woodpile.m:165-182),
    closing Gamma dropped: 4 segments x part_num points = 40 k for part_num = 10."""
    G = np.zeros(3)
    X = np.array([np.pi, 0, 0])
    M = np.array([np.pi, np.pi, 0])
    R = np.array([np.pi, np.pi, np.pi])
    ks = []
    pts = [G, X, M, R, G]
    for i in range(4):
        seg = np.linspace(pts[i], pts[i + 1], part_num + 1)
        ks.extend(seg[:-1])
    return np.array(ks)


def fcc_band_path(n_k, part_num=None):
    """n_k points along X-U-L-Gamma-X-W-K of the FCC zone (reference path string 'XULGXWK')."""
    B = 2 * np.pi * np.linalg.inv(FCC_LATTICE.T)
    v = {"G": [0, 0, 0], "K": [3 / 8, 3 / 8, 3 / 4], "L": [0.5, 0.5, 0.5], "U": [5 / 8, 1 / 4, 5 / 8],
         "W": [0.5, 1 / 4, 3 / 4], "X": [0.5, 0, 0.5]}
    names = "XULGXWK"
    pts = [B @ np.array(v[c], float) for c in names]
    seglen = np.array([np.linalg.norm(pts[i + 1] - pts[i]) for i in range(len(pts) - 1)])
    t = np.linspace(0, seglen.sum(), n_k, endpoint=False)
    out = []
    cum = np.concatenate([[0], np.cumsum(seglen)])
    for tt in t:
        i = min(np.searchsorted(cum, tt, side="right") - 1, len(seglen) - 1)
        f = (tt - cum[i]) / seglen[i]
        out.append(pts[i] + f * (pts[i + 1] - pts[i]))
    return np.array(out)


# ---------------------------------------------------------------------------
# FCC skewed-grid index sets (light_locked_Huang/Construct_idx_all_FCC.m)
# ---------------------------------------------------------------------------

def _fcc_sites(edge_len):
    """S1..S5 of Construct_idx_all_FCC.m:150-155 (origin, a1, a2, a3, and the interior site)."""
    r3, r6 = np.sqrt(3.0), np.sqrt(6.0)
    return edge_len * np.array([[0, 0, 0], [1, 0, 0], [0.5, r3 / 2, 0], [0.5, r3 / 6, r6 / 3],
                                [0.5, r3 / 6, r6 / 12]], float)


def _fcc_locate(xyz, shape, rd_sphere, rd_cylinder, edge_len):
    S = _fcc_sites(edge_len)
    inside = np.zeros(xyz.shape[0], bool)
    if shape == "cylinder":  # locate_cyld: rods S5-S_i (radius rd_cylinder, between the two sites)
        for i in range(4):
            d = S[i] - S[4]
            L2 = d @ d
            u = d / np.sqrt(L2)
            proj = (xyz - S[4]) @ np.outer(u, u).T + S[4]
            D1 = np.sum((xyz - proj) ** 2, axis=1)
            D2 = np.sum((proj - S[i]) ** 2, axis=1)
            D3 = np.sum((proj - S[4]) ** 2, axis=1)
            inside |= (D1 <= rd_cylinder ** 2) & (D2 <= L2) & (D3 <= L2)
    elif shape == "ellipse":  # locate_ellipse: prolate spheroids with foci S_i, S5
        d5 = np.linalg.norm(xyz - S[4], axis=1)
        for i in range(4):
            s = np.linalg.norm(xyz - S[i], axis=1) + d5
            inside |= s <= 2 * np.sqrt(rd_cylinder ** 2 + (np.linalg.norm(S[i] - S[4]) / 2) ** 2)
    else:
        raise ValueError("shape must be 'cylinder' or 'ellipse'")
    for i in range(5):  # spheres at the five sites
        inside |= np.sum((xyz - S[i]) ** 2, axis=1) <= rd_sphere ** 2
    return inside


def fcc_index_sets(dims, edge_len, rd_sphere, rd_cylinder, delta, shape="cylinder"):
    """Inside masks of Construct_idx_all_FCC.m: grid points of each Yee location (edges X/Y/Z,
    cube centres, faces X/Y/Z, vertices), half-offset by 0.5 mesh along the named axes, x fastest.
    Returns (edge_sets, face_sets) = ([X, Y, Z, Cubic], [X, Y, Z, vertex]) boolean masks, the
    inputs of mtx_Eps_Mu_Zeta_Xi (edge: edge X/Y/Z + Cubic; face: face X/Y/Z + vertex)."""
    n1, n2, n3 = dims

    def pts(ox, oy, oz):
        i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
        return np.stack([(i1.ravel() + ox) * delta[0], (i2.ravel() + oy) * delta[1], (i3.ravel() + oz) * delta[2]], 1)

    loc = lambda o: _fcc_locate(pts(*o), shape, rd_sphere, rd_cylinder, edge_len)  # noqa: E731
    edge = [loc((0.5, 0, 0)), loc((0, 0.5, 0)), loc((0, 0, 0.5)), loc((0.5, 0.5, 0.5))]
    face = [loc((0, 0.5, 0.5)), loc((0.5, 0, 0.5)), loc((0.5, 0.5, 0)), loc((0, 0, 0))]
    return edge, face

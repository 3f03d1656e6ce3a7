"""Round trips of the reference output formats (CPU): the band-structure .mat file read
back with scipy.io.loadmat, and the JD-SIRA log parsed with the pattern the reference's
figure script uses (draw_test/Fig6ab_Err_vs_Iternum_Data_Extraction.m:10)."""
import re

import numpy as np
import scipy.io


def test_band_structure_mat_round_trip(pkg, tmp_path):
    rng = np.random.default_rng(1)
    nk, nev, N = 3, 4, 27
    F = rng.uniform(0.1, 1, (nev, nk))
    e_h = rng.standard_normal((nk, nev, 6 * N)) + 1j * rng.standard_normal((nk, nev, 6 * N))
    rn, ro = rng.uniform(0, 1e-9, (nk, nev)), rng.uniform(0, 1e-9, (nk, nev))
    path = tmp_path / "band.mat"
    pkg.outputs.save_band_structure(str(path), F, e_h, rn, ro, [1.0, 2.0, 3.0], [100, 120, 90])
    m = scipy.io.loadmat(str(path))
    assert np.array_equal(m["Freq_mtx"], F)
    assert m["iter_records"].shape == (nk, 1) and m["iter_records"][1, 0] == 120
    assert m["result"].shape == (1, nk)
    for i in range(nk):
        r = m["result"][0, i]
        em = r["Eigenmode"][0, 0]
        for c, name in enumerate(("e1", "e2", "e3", "h1", "h2", "h3")):
            assert np.array_equal(em[name][0, 0], e_h[i][:, c * N:(c + 1) * N].T)
        assert np.array_equal(r["res_NFSEP"][0, 0][:, 0], rn[i])
    assert np.array_equal(m["time_minres"]["compute_eigs"][0, 0][0], [1.0, 2.0, 3.0])


def test_jdsira_log_parses_like_the_reference(pkg):
    class R:
        pass

    res = R()
    res.log = [(1, 1, 1.25, 0.3), (1, 2, 1.2001 + 1e-9j, 2e-3), (1, 3, 1.2, 5e-10), (2, 4, 1.5 - 1e-9j, 1e-2),
               (2, 5, 1.49, 8e-10)]
    res.locked = [(1, 4, 0.5, 1.2, 5e-10), (2, 6, 0.9, 1.49, 8e-10)]
    res.inner_per_lock = [100, 80]
    lines = pkg.outputs.jdsira_log_lines(res)
    text = "\n".join(lines)
    pattern = r'rsdl\(\s*(\d+)\s*,\s*(\d+)\s*\)\s*=\s*([+-]?\d*\.?\d+(?:[eE][+-]?\d+)?)'
    toks = re.findall(pattern, text)
    assert [(int(a), int(b)) for a, b, _ in toks] == [(o, it) for (o, it, _, _) in res.log]
    assert np.allclose([float(v) for _, _, v in toks], [r for (_, _, _, r) in res.log], rtol=1e-2)
    assert "ew_crnt( 1,   2) = 1.200100000000e+00+1.000000000000e-09i;" in text
    assert sum(1 for ln in lines if ln.startswith("itno(kk,")) == 2
    assert "total_inner_iter(kk, 2) =      80; " in lines
    assert lines[-2] == "kk = kk +1; "

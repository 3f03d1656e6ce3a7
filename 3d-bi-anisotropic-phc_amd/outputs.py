"""The reference's output formats, so its figure scripts can read this library's results.

* ``save_band_structure``  the Lanczos band-path file,
  ``save(filename, 'Freq_mtx', 'result', 'time_minres', 'iter_records')``
  (band_structure_woodpile_Lanczos_n60_gamma0p2.m:122-123): Freq_mtx (nev x nk),
  result{i}.Eigenmode.{e1,e2,e3,h1,h2,h3} (N x nev each), result{i}.res_NFSEP /
  res_ORGSEP (nev x 1), time_minres.compute_eigs (1 x nk), iter_records (nk x 1).
* ``save_jdsira``          the JD-SIRA file ``save(filename, 'ew', 'result', 'wave_vec',
  'time_minres')`` (light_locked_Huang/GEP_AB_Herm_JDSIRA_Driver4.m:69-76).
* ``jdsira_log_lines``     the text log of GEP_AB_Herm_JDSIRA_Locking4.m:184-195 (one
  ``ew_crnt(i, j) = ...; rsdl(i, j) = ...;`` line per outer step), output.m:3-12 (one
  ``itno(kk, i) = ...`` line per locked pair), Driver4.m:65 (``total_inner_iter``) and
  Driver4.m:79 (``kk = kk +1;``), which draw_test/Fig6ab_Err_vs_Iternum_Data_Extraction.m
  parses.
"""
from __future__ import annotations

import numpy as np
import scipy.io


def _eigenmode(e_h, N):
    """[e; h] rows (nvec, 6N) -> dict of the six N x nvec component blocks."""
    e_h = np.asarray(e_h)
    out = {}
    for c, name in enumerate(("e1", "e2", "e3", "h1", "h2", "h3")):
        out[name] = np.ascontiguousarray(e_h[:, c * N:(c + 1) * N].T)
    return out


def save_band_structure(path, Freq_mtx, e_h, res_NFSEP, res_ORGSEP, compute_eigs, iter_records, opt=None):
    """``Freq_mtx`` (nev, nk); ``e_h`` (nk, nev, 6N) restored [e; h]; residuals (nk, nev);
    ``compute_eigs`` (nk,) seconds; ``iter_records`` (nk,) Lanczos iterations."""
    Freq_mtx = np.asarray(Freq_mtx)
    nk = Freq_mtx.shape[1]
    N = np.asarray(e_h).shape[2] // 6
    result = np.empty((1, nk), dtype=object)
    for i in range(nk):
        r = {"Eigenmode": _eigenmode(e_h[i], N),
             "res_NFSEP": np.asarray(res_NFSEP[i], float).reshape(-1, 1),
             "res_ORGSEP": np.asarray(res_ORGSEP[i], float).reshape(-1, 1)}
        if opt is not None:
            r["opt"] = opt
        result[0, i] = r
    scipy.io.savemat(path, {"Freq_mtx": Freq_mtx, "result": result,
                            "time_minres": {"compute_eigs": np.asarray(compute_eigs, float).reshape(1, -1)},
                            "iter_records": np.asarray(iter_records, float).reshape(-1, 1)})


def save_jdsira(path, ew, e_h, res_NFSEP, res_ORGSEP, wave_vec, seconds):
    """JD-SIRA results: ``ew`` (nev,), ``e_h`` (nev, 6N), residuals (nev,), ``wave_vec`` (3,)."""
    N = np.asarray(e_h).shape[1] // 6
    result = {"Eigenmode": _eigenmode(e_h, N), "res_NFSEP": np.asarray(res_NFSEP, float).reshape(-1, 1),
              "res_ORGSEP": np.asarray(res_ORGSEP, float).reshape(-1, 1)}
    scipy.io.savemat(path, {"ew": np.asarray(ew).reshape(-1, 1), "result": result,
                            "wave_vec": np.asarray(wave_vec, float).reshape(-1, 1),
                            "time_minres": {"compute_an_eig": float(seconds)}})


def _cplx(z, fmt):
    z = complex(z)
    if z.imag > 0:
        return (fmt + "+" + fmt + "i") % (z.real, z.imag)
    if z.imag < 0:
        return (fmt + "-" + fmt + "i") % (z.real, -z.imag)
    return fmt % z.real


def jdsira_log_lines(res):
    """Log lines of one JD-SIRA run (a JDResult) in the reference's format."""
    lines = []
    locks = list(res.locked)
    inner = list(getattr(res, "inner_per_lock", []))
    li = 0
    log = list(res.log)
    for n, (order, it, ritz, rsdl) in enumerate(log):
        lines.append("ew_crnt(%2.0f,%4.0f) = %s; rsdl(%2.0f,%4.0f) = %8.2e;" % (order, it, _cplx(ritz, "%18.12e"), order,
                                                                                it, rsdl))
        last_of_call = n + 1 == len(log) or log[n + 1][0] != order
        while last_of_call and li < len(locks) and locks[li][0] == order:
            o, e, secs, ew, r = locks[li]
            lines.append("itno(kk,%3.0f) = %4.0f; CPUtime(kk,%3.0f) = %8.1f; ew(kk,%3.0f) = %s; rsdl(kk,%3.0f) = %13.4e;"
                         % (o, e, o, secs, o, _cplx(ew, "%24.16e"), o, r))
            if li < len(inner):
                lines.append("total_inner_iter(kk,%2.0f) = %7.0f; " % (o, inner[li]))
            li += 1
    lines.append("kk = kk +1; ")
    lines.append("")
    return lines


def write_jdsira_log(path, res):
    with open(path, "w") as f:
        f.write("\n".join(jdsira_log_lines(res)) + "\n")

"""Python mirror of the reference's operator interface over the C-ABI (``libfame.so``).

The reference is MATLAB; its eigensolvers see the hot path through function
handles with these signatures (names kept):

* ``MtxVecMul_FFT(vec_in, plan, istrans)``  subroutines/gMtxVecMul_FFT.m:1
* ``fnc.Pr / Qr / Prs / Qrs``                fun_Lanczos/Funhand_Generate_Lanzcos.m:79-82
* ``fncA(x, plan)``                          fun_Lanczos/fncA.m:1
* ``fncB(x, plan)``                          fun_Lanczos/fncB.m:1
* ``fncinvB(x, plan, tol, maxit)``           fun_Lanczos/fncinvB.m:1
* ``Lanczos(plan, k, tol, maxiter, p, v0)``  fun_Lanczos/Lanczos.m:1
* ``Eigen_Restoration_Chiral_no_inv``        fun_no_inv/Eigen_Restoration_Chiral_no_inv.m:1

Every call runs the HIP kernels of ``libfame.so`` on the GPU.  There is no CPU
fallback: if the library is missing or no GPU is present the calls raise.

Vectors are complex128 torch tensors on a CUDA (HIP) device, shaped
``(nb, len)`` (or ``(len,)`` when ``nb == 1``), in the reference's MATLAB
column-major order.  The plan runs on torch's current stream, so results are
ordered with surrounding torch work without extra synchronisation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfame.so")

FAME_OK = 0
FAME_ERR_ARG = -1
FAME_ERR_HIP = -2
FAME_ERR_UNSUPPORTED = -3
FAME_ERR_NOT_CONVERGED = -4

_lib = None


class FameError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fame error {code}: {msg}")
        self.code = code


class NotConverged(FameError):
    pass


def load_library(path: str = LIB_PATH):
    """Load ``libfame.so`` and declare the C-ABI signatures (include/fame_abi.h)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FameError(FAME_ERR_HIP, f"{path} is not built; run build() first (no CPU fallback exists)")
    lib = ctypes.CDLL(path)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    pd, pi_ = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    sig = {
        "fame_version": ([], ctypes.c_char_p),
        "fame_last_error": ([], ctypes.c_char_p),
        "fame_plan_create": ([ctypes.POINTER(vp), i, i, i, i, pd, pd, i, pd, ctypes.c_void_p], i),
        "fame_plan_set_theta": ([vp, pd], i),
        "fame_plan_set_structure": ([vp, i, pd, pd], i),
        "fame_plan_material_ids": ([vp, vp], i),
        "fame_plan_destroy": ([vp], i),
        "fame_sync": ([vp], i),
        "fame_plan_stream": ([vp], vp),
        "fame_plan_set_stream": ([vp, vp], i),
        "fame_MtxVecMul_FFT": ([vp, i, vp, vp], i),
        "fame_Pr": ([vp, vp, vp], i),
        "fame_Qr": ([vp, vp, vp], i),
        "fame_Prs": ([vp, vp, vp], i),
        "fame_Qrs": ([vp, vp, vp], i),
        "fame_fncA": ([vp, vp, vp], i),
        "fame_fncB": ([vp, vp, vp], i),
        "fame_material_apply": ([vp, vp, vp], i),
        "fame_fncinvB": ([vp, vp, vp, d, i, pi_, pd], i),
        "fame_lanczos": ([vp, i, d, i, i, vp, d, i, vp, pd, pi_], i),
        "fame_eigen_restoration": ([vp, i, vp, vp], i),
        "fame_dot": ([vp, vp, vp, ctypes.c_longlong, vp], i),
        "fame_minres": ([vp, vp, vp, d, d, i, pi_, pd], i),
        "fame_block_dot": ([vp, vp, i, vp, ctypes.c_longlong, vp], i),
        "fame_block_gemv": ([vp, vp, i, ctypes.c_longlong, vp, d, vp], i),
        "fame_block_dot_noalloc": ([vp, vp, i, vp, ctypes.c_longlong, vp], i),
        "fame_mgs": ([vp, vp, i, ctypes.c_longlong, vp, vp], i),
        "fame_jd_ritz_residual": ([vp, vp, i, ctypes.c_longlong, pd, d, d, vp, vp, pd], i),
        "fame_jd_expand": ([vp, vp, i, ctypes.c_longlong, vp, d, d, i, pi_, pd], i),
        "fame_plan_profile": ([vp, i], i),
        "fame_plan_set_generic": ([vp, i], i),
        "fame_plan_set_cg_fused": ([vp, i], i),
        "fame_plan_set_cg_graph": ([vp, i], i),
        "fame_plan_set_streams": ([vp, i], i),
        "fame_plan_get_streams": ([vp, pi_], i),
        "fame_plan_set_material_tile": ([vp, i, i, i], i),
        "fame_plan_get_material_tile": ([vp, pi_, pi_, pi_], i),
        "fame_plan_set_chunk": ([vp, i], i),
        "fame_plan_get_chunk": ([vp], i),
        "fame_plan_autotune": ([vp, pd], i),
        "fame_plan_set_variant": ([vp, i, i], i),
        "fame_plan_get_variant": ([vp, pi_, pi_], i),
        "fame_plan_stage_times": ([vp, pd, ctypes.POINTER(ctypes.c_longlong)], i),
        "fame_plan_material_kernel_time": ([vp, pd, ctypes.POINTER(ctypes.c_longlong)], i),
        "fame_plan_set_vertex_material": ([vp, pd, pd, d, pd], i),
        "fame_material_gradient": ([vp, i, vp, vp], i),
        "fame_stream_probe": ([i, vp, vp, ctypes.c_longlong, i, vp, pd], i),
        "fame_fcc_create": ([ctypes.POINTER(vp), i, i, i, pd, pd, pd, pd, pd], i),
        "fame_fcc_destroy": ([vp], i),
        "fame_fcc_set_stream": ([vp, vp], i),
        "fame_fcc_sync": ([vp], i),
        "fame_fcc_set_generic": ([vp, i], i),
        "fame_fcc_set_nfgep_variant": ([vp, i], i),
        "fame_fcc_T": ([vp, i, vp, vp], i),
        "fame_fcc_TH": ([vp, i, vp, vp], i),
        "fame_fcc_Sigma_r": ([vp, vp], i),
        "fame_fcc_Pr": ([vp, vp, vp], i),
        "fame_fcc_Qr": ([vp, vp, vp], i),
        "fame_fcc_Prs": ([vp, vp, vp], i),
        "fame_fcc_Qrs": ([vp, vp, vp], i),
        "fame_fcc_nfgep_A": ([vp, vp, vp, vp, vp, vp], i),
        "fame_fcc_eigenvector": ([vp, vp, vp, vp, vp, vp], i),
        "fame_fcc_nfgep_A_ex": ([vp, vp, vp, vp, vp, vp, i], i),
        "fame_fcc_eigenvector_ex": ([vp, vp, vp, vp, vp, vp, i], i),
        "fame_fcc_nfgep_data": ([vp, vp, vp, vp], i),
        "fame_fcc_set_material": ([vp, vp, vp, pd], i),
        "fame_fcc_param_apply": ([vp, i, vp, vp], i),
        "fame_fcc_gep_B": ([vp, vp, vp], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def exported_symbols():
    """Names of the C-ABI entry points this wrapper binds."""
    return [
        "fame_version", "fame_last_error", "fame_plan_create", "fame_plan_set_theta", "fame_plan_destroy",
        "fame_plan_set_structure", "fame_plan_material_ids",
        "fame_sync", "fame_plan_stream", "fame_plan_set_stream", "fame_MtxVecMul_FFT", "fame_Pr", "fame_Qr",
        "fame_Prs", "fame_Qrs", "fame_fncA", "fame_fncB", "fame_material_apply", "fame_fncinvB", "fame_lanczos",
        "fame_eigen_restoration", "fame_dot", "fame_minres", "fame_block_dot", "fame_block_gemv", "fame_mgs", "fame_block_dot_noalloc", "fame_jd_ritz_residual", "fame_jd_expand",
        "fame_plan_profile", "fame_plan_stage_times", "fame_plan_material_kernel_time", "fame_plan_set_cg_fused", "fame_plan_set_cg_graph",
        "fame_plan_set_streams", "fame_plan_get_streams",
        "fame_plan_set_material_tile", "fame_plan_get_material_tile",
        "fame_plan_set_generic", "fame_plan_set_chunk", "fame_plan_get_chunk", "fame_plan_autotune",
        "fame_plan_set_variant", "fame_plan_get_variant",
        "fame_plan_set_vertex_material", "fame_material_gradient", "fame_stream_probe",
        "fame_fcc_create", "fame_fcc_destroy", "fame_fcc_set_stream", "fame_fcc_sync", "fame_fcc_set_generic", "fame_fcc_set_nfgep_variant", "fame_fcc_T", "fame_fcc_TH",
        "fame_fcc_Sigma_r", "fame_fcc_Pr", "fame_fcc_Qr", "fame_fcc_Prs", "fame_fcc_Qrs", "fame_fcc_nfgep_A",
        "fame_fcc_eigenvector", "fame_fcc_nfgep_A_ex", "fame_fcc_eigenvector_ex", "fame_fcc_nfgep_data",
        "fame_fcc_set_material", "fame_fcc_param_apply", "fame_fcc_gep_B",
    ]


def _check(rc):
    if rc == FAME_OK:
        return
    msg = _lib.fame_last_error().decode()
    if rc == FAME_ERR_NOT_CONVERGED:
        raise NotConverged(rc, msg)
    raise FameError(rc, msg)


def _dptr(arr):
    """Pointer to a float64 C-contiguous array; the returned ctypes object keeps ``arr`` alive,
    so ``arr`` must already be float64 and contiguous (no temporary copy is made here)."""
    if arr.dtype != np.float64 or not arr.flags.c_contiguous:
        raise TypeError("_dptr needs a C-contiguous float64 array")
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def stream_probe(kind, src=None, dst=None, reps=10, stream=None):
    """Mean milliseconds per pass of the HBM stream probe (fame_stream_probe): kind "read" (src),
    "write" (dst) or "copy" (src -> dst); src / dst are contiguous device tensors of equal byte
    size, a multiple of 16 bytes."""
    lib = load_library()
    k = {"read": 0, "write": 1, "copy": 2}[kind]
    buf = src if src is not None else dst
    nbytes = buf.numel() * buf.element_size()
    ms = ctypes.c_double()
    ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _check(lib.fame_stream_probe(k, ptr(src), ptr(dst), nbytes, int(reps), stream, ctypes.byref(ms)))
    return ms.value


class Plan:
    """One grid + material with ``nb`` Bloch vectors; owns the device tables.

    Parameters mirror what the reference's per-k setup consumes:
    ``theta`` (nb,3) = k . a_d, ``delta`` (3,) mesh lengths, ``mat_table``
    (n_mat, 3, 3, 3) complex = per material the N-, M-, J-slot tensors, and
    ``mat_id`` (N,) uint8 per-vertex material index (x fastest).
    """

    def __init__(self, dims, theta, mat_table, mat_id=None, delta=None, device=None):
        import torch

        self.lib = load_library()
        if not torch.cuda.is_available():
            raise FameError(FAME_ERR_HIP, "no GPU visible: the HIP path has no CPU fallback")
        self.torch = torch
        self.device = torch.device("cuda" if device is None else device)
        self.dims = tuple(int(x) for x in dims)
        self.N = int(np.prod(self.dims))
        th = np.ascontiguousarray(np.atleast_2d(np.asarray(theta, dtype=np.float64)))
        self.nb = th.shape[0]
        self.theta = th.copy()
        self.delta = np.ascontiguousarray(
            np.array([1.0 / d for d in self.dims]) if delta is None else np.asarray(delta, np.float64))
        tab = np.ascontiguousarray(np.asarray(mat_table, dtype=np.complex128).reshape(-1, 27))
        self.n_mat = tab.shape[0]
        self._tab = tab
        ids = None
        if mat_id is not None:
            ids = np.ascontiguousarray(np.asarray(mat_id, dtype=np.uint8).reshape(-1))
            if ids.size != self.N:
                raise ValueError("mat_id must have N entries")
        self._ids = ids
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            rc = self.lib.fame_plan_create(
                ctypes.byref(h), self.dims[0], self.dims[1], self.dims[2], self.nb, _dptr(th), _dptr(self.delta),
                self.n_mat, ctypes.cast(tab.ctypes.data, ctypes.POINTER(ctypes.c_double)),
                None if ids is None else ids.ctypes.data)
        _check(rc)
        self.h = h
        _check(self.lib.fame_plan_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)))

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.fame_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_theta(self, theta):
        th = np.ascontiguousarray(np.atleast_2d(np.asarray(theta, dtype=np.float64)))
        if th.shape != (self.nb, 3):
            raise ValueError("theta must be (nb, 3)")
        self.theta = th.copy()
        _check(self.lib.fame_plan_set_theta(self.h, _dptr(th)))

    STRUCTURES = ("woodpile", "sphere", "tetra")

    def set_structure(self, kind, size, center=(0.5, 0.5, 0.5), lattice=None, a=1.0):
        """Per-vertex material ids computed on the device (fame_plan_set_structure): "woodpile"
        (rod width ``size``, lattice constant ``a``), "sphere" (radius ``size``) or "tetra" (edge
        ``size``) centred at lattice coordinates ``center``, lattice vectors as the columns of
        ``lattice``; id 1 inside, 0 outside."""
        k = self.STRUCTURES.index(kind)
        params = np.ascontiguousarray([float(size), float(a), 0.0, 0.0] if k == 0 else
                                      [float(size), *map(float, center)], dtype=np.float64)
        lat = np.ascontiguousarray(np.eye(3) if lattice is None else np.asarray(lattice, float)).reshape(-1)
        _check(self.lib.fame_plan_set_structure(self.h, k, _dptr(params), _dptr(lat)))

    def material_ids(self):
        """The per-vertex material ids the stencil reads (device uint8 tensor, N)."""
        out = self.torch.empty(self.N, dtype=self.torch.uint8, device=self.device)
        _check(self.lib.fame_plan_material_ids(self.h, ctypes.c_void_p(out.data_ptr())))
        return out

    def sync(self):
        _check(self.lib.fame_sync(self.h))

    def set_chunk(self, chunk):
        """fncB processes the k-batch in chunks of ``chunk`` items (0 = all)."""
        _check(self.lib.fame_plan_set_chunk(self.h, int(chunk)))

    def autotune(self):
        """Pick the fastest FFT launch shapes and fncB chunk size on this plan;
        returns (chunk, ms per fncB)."""
        ms = ctypes.c_double()
        _check(self.lib.fame_plan_autotune(self.h, ctypes.byref(ms)))
        return self.lib.fame_plan_get_chunk(self.h), ms.value

    def set_variant(self, xy_variant, z_variant):
        _check(self.lib.fame_plan_set_variant(self.h, int(xy_variant), int(z_variant)))

    def variant(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.fame_plan_get_variant(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_generic(self, generic=True):
        """Force the runtime-radix FFT kernels (True) or allow the specialised ones (False)."""
        _check(self.lib.fame_plan_set_generic(self.h, int(bool(generic))))

    STAGES = ("zpass_mix_in", "xypass_inv", "material", "xypass_fwd", "zpass_mix_out")

    def set_cg_graph(self, on=True):
        """Replay the fused CG iteration from a captured HIP graph, or launch one by one (default)."""
        _check(self.lib.fame_plan_set_cg_graph(self.h, int(bool(on))))

    def set_streams(self, nstreams):
        """Queue fncB / lock-step CG chunks on one stream or deal them over 2 or 3 streams."""
        _check(self.lib.fame_plan_set_streams(self.h, int(nstreams)))

    def streams(self):
        n = ctypes.c_int()
        _check(self.lib.fame_plan_get_streams(self.h, ctypes.byref(n)))
        return n.value

    def set_material_tile(self, tx, ty, tz=0):
        """Material-stencil tile (tx, ty columns, tz slices; 0, 0 = heuristic)."""
        _check(self.lib.fame_plan_set_material_tile(self.h, int(tx), int(ty), int(tz)))

    def material_tile(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(self.lib.fame_plan_get_material_tile(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_cg_fused(self, fused=True, seq_chunk=0):
        """CG schedule for fncinvB and Lanczos (same iterates up to rounding): fused lock step
        (default), fused with chunks of ``seq_chunk`` k-points solved one after another, or the
        textbook loop (fused=False)."""
        v = 0 if not fused else (1 + int(seq_chunk) if seq_chunk > 0 else 1)
        _check(self.lib.fame_plan_set_cg_fused(self.h, v))

    def profile(self, enable=True):
        """Per-kernel HIP-event timing of fncB (measurement aid)."""
        _check(self.lib.fame_plan_profile(self.h, int(bool(enable))))

    def stage_times(self):
        ms = (ctypes.c_double * 5)()
        calls = ctypes.c_longlong()
        _check(self.lib.fame_plan_stage_times(self.h, ms, ctypes.byref(calls)))
        return dict(zip(self.STAGES, ms[:])), calls.value

    def material_kernel_time(self):
        """(ms, launches): the material stencil's own kernel time while profiling (its launch
        events), without the gaps the stage events include."""
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        _check(self.lib.fame_plan_material_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # -- buffers ------------------------------------------------------------------
    def empty(self, ncomp):
        return self.torch.empty((self.nb, ncomp * self.N), dtype=self.torch.complex128, device=self.device)

    def _as_dev(self, x, ncomp):
        t = self.torch
        if not isinstance(x, t.Tensor):
            x = t.as_tensor(np.asarray(x, dtype=np.complex128))
        x = x.to(device=self.device, dtype=t.complex128).contiguous()
        if x.numel() != self.nb * ncomp * self.N:
            raise ValueError(f"expected {self.nb} x {ncomp}N complex values, got {x.numel()}")
        return x.reshape(self.nb, ncomp * self.N)

    def _shape_out(self, y, like):
        return y.reshape(like.shape) if like.dim() == 1 or like.shape == y.shape else y

    def _call(self, fn, x, nin, nout, *extra):
        xin = self._as_dev(x, nin)
        y = self.empty(nout)
        _check(fn(self.h, *extra, ctypes.c_void_p(xin.data_ptr()), ctypes.c_void_p(y.data_ptr())))
        return y.reshape(-1) if self.nb == 1 else y

    # -- operators (MATLAB names) ------------------------------------------------
    def MtxVecMul_FFT(self, vec_in, istrans):
        """Psi/Phi ('notranspose') or Psis/Phis ('transpose') on [nb][3N] (gMtxVecMul_FFT.m:1-61)."""
        flag = {"notranspose": 0, "transpose": 1}[istrans]
        return self._call(self.lib.fame_MtxVecMul_FFT, vec_in, 3, 3, flag)

    def Pr(self, x):
        return self._call(self.lib.fame_Pr, x, 2, 3)

    def Qr(self, x):
        return self._call(self.lib.fame_Qr, x, 2, 3)

    def Prs(self, y):
        return self._call(self.lib.fame_Prs, y, 3, 2)

    def Qrs(self, y):
        return self._call(self.lib.fame_Qrs, y, 3, 2)

    def fncA(self, x):
        return self._call(self.lib.fame_fncA, x, 4, 4)

    def fncB(self, x):
        return self._call(self.lib.fame_fncB, x, 4, 4)

    def material_apply(self, he):
        """[M h + J^H e; J h + N e] on twisted real-space [h; e] (fncB.m:4-5)."""
        return self._call(self.lib.fame_material_apply, he, 6, 6)

    def fncinvB(self, x, tol=1e-12, maxit=10000, raise_on_maxit=True):
        """CG solve of B y = x (fncinvB.m:4); returns (y, iters, relres)."""
        xin = self._as_dev(x, 4)
        y = self.empty(4)
        iters = (ctypes.c_int * self.nb)()
        relres = (ctypes.c_double * self.nb)()
        rc = self.lib.fame_fncinvB(self.h, ctypes.c_void_p(xin.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                   float(tol), int(maxit), iters, relres)
        if rc != FAME_ERR_NOT_CONVERGED or raise_on_maxit:
            _check(rc)
        y = y.reshape(-1) if self.nb == 1 else y
        return y, np.array(iters[:]), np.array(relres[:])

    def minres(self, x, sigma=0.0, tol=1e-2, maxit=30000, raise_on_maxit=False):
        """MINRES on (fncB - sigma fncA) y = x (the SIRA inner solve); returns (y, iters, relres)."""
        xin = self._as_dev(x, 4)
        y = self.empty(4)
        iters = (ctypes.c_int * self.nb)()
        relres = (ctypes.c_double * self.nb)()
        rc = self.lib.fame_minres(self.h, ctypes.c_void_p(xin.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                  float(sigma), float(tol), int(maxit), iters, relres)
        if rc != FAME_ERR_NOT_CONVERGED or raise_on_maxit:
            _check(rc)
        y = y.reshape(-1) if self.nb == 1 else y
        return y, np.array(iters[:]), np.array(relres[:])

    def block_dot(self, V, w):
        """V^H w for V (m, len) and w (len,), device tensors; returns a device tensor (m,)."""
        m, L = V.shape
        out = self.torch.empty(m, dtype=self.torch.complex128, device=self.device)
        _check(self.lib.fame_block_dot(self.h, ctypes.c_void_p(V.data_ptr()), int(m), ctypes.c_void_p(w.data_ptr()),
                                       int(L), ctypes.c_void_p(out.data_ptr())))
        return out

    def block_gemv(self, V, y, out=None, beta=0.0):
        """out = beta*out + V^T-combination sum_j y_j V_j for V (m, len); y device (m,)."""
        m, L = V.shape
        if out is None:
            out = self.torch.empty(L, dtype=self.torch.complex128, device=self.device)
            beta = 0.0
        y = y.to(device=self.device, dtype=self.torch.complex128).contiguous()
        _check(self.lib.fame_block_gemv(self.h, ctypes.c_void_p(V.data_ptr()), int(m), int(L),
                                        ctypes.c_void_p(y.data_ptr()), float(beta), ctypes.c_void_p(out.data_ptr())))
        return out

    def mgs(self, V, t, coef=False):
        """Modified Gram-Schmidt of t against the rows of V, in order, in place on the device
        (GEP_AB_Herm_JDSIRA_Locking4.m:326-330); returns the coefficients V_j^H t (device) if
        ``coef``."""
        m, L = V.shape
        if t.numel() != L or not t.is_contiguous() or not V.is_contiguous():
            raise ValueError("mgs: V (m, len) and t (len,) must be contiguous device tensors")
        c = self.torch.empty(m, dtype=self.torch.complex128, device=self.device) if coef else None
        _check(self.lib.fame_mgs(self.h, ctypes.c_void_p(V.data_ptr()), int(m), int(L), ctypes.c_void_p(t.data_ptr()),
                                 None if c is None else ctypes.c_void_p(c.data_ptr())))
        return c

    def jd_ritz_residual(self, V, y, theta):
        """u = V^T y / ||.||, r = fncB u - theta fncA u and ||r|| for V (m, len) device rows and host
        coefficients y (m,) (Locking4.m:170-178); returns (u, r, ||r||)."""
        m, L = V.shape
        yv = np.ascontiguousarray(np.asarray(y, complex).reshape(m)).view(np.float64)
        u = self.torch.empty(L, dtype=self.torch.complex128, device=self.device)
        r = self.torch.empty_like(u)
        rn = ctypes.c_double()
        th = complex(theta)
        _check(self.lib.fame_jd_ritz_residual(self.h, ctypes.c_void_p(V.data_ptr()), int(m), int(L), _dptr(yv),
                                              th.real, th.imag, ctypes.c_void_p(u.data_ptr()),
                                              ctypes.c_void_p(r.data_ptr()), ctypes.byref(rn)))
        return u, r, rn.value

    def jd_expand(self, V, m, r, sigma, tol, maxit):
        """MINRES correction of r, orthogonalised and normalised into row m of V (a contiguous
        (rows, len) device tensor), and the new column of V^H fncB V and V^H fncA V over rows
        0..m (Locking4.m:311-341); returns (MINRES steps, column of V^H A V, column of V^H B V)."""
        L = V.shape[1]
        if not V.is_contiguous() or m + 1 > V.shape[0]:
            raise ValueError("jd_expand: V must be contiguous with room for row m")
        proj = np.zeros(2 * (m + 1), complex)
        its = ctypes.c_int()
        pv = proj.view(np.float64)
        _check(self.lib.fame_jd_expand(self.h, ctypes.c_void_p(V.data_ptr()), int(m), int(L),
                                       ctypes.c_void_p(r.data_ptr()), float(sigma), float(tol), int(maxit),
                                       ctypes.byref(its), _dptr(pv)))
        return its.value, proj[:m + 1], proj[m + 1:]

    def Lanczos(self, k, tol=1e-10, maxiter=10000, p=None, v0=None, cg_tol=1e-12, cg_maxit=10000,
                raise_on_maxit=True):
        """Thick-restart Lanczos (Lanczos.m:1-156); returns (ev [nb,k,4N], ew [nb,k], iters [nb]).

        iters[b] < 0 marks a k-point that did not converge (its ev/ew are the last Ritz pairs);
        with raise_on_maxit (default) any such k-point raises NotConverged instead."""
        p = 2 * k if p is None else p
        if v0 is None:
            v0 = self.torch.ones((self.nb, 4 * self.N), dtype=self.torch.complex128, device=self.device)
        v = self._as_dev(v0, 4)
        ev = self.torch.empty((self.nb, k, 4 * self.N), dtype=self.torch.complex128, device=self.device)
        ew = np.zeros((self.nb, k))
        iters = (ctypes.c_int * self.nb)()
        rc = self.lib.fame_lanczos(self.h, int(k), float(tol), int(maxiter), int(p), ctypes.c_void_p(v.data_ptr()),
                                   float(cg_tol), int(cg_maxit), ctypes.c_void_p(ev.data_ptr()),
                                   ew.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), iters)
        if rc != FAME_ERR_NOT_CONVERGED or raise_on_maxit:
            _check(rc)
        return ev, ew, np.array(iters[:])

    def Eigen_Restoration_Chiral_no_inv(self, ev):
        """(e, h) fields from NFSEP eigenvectors (Eigen_Restoration_Chiral_no_inv.m:3-25).

        ``ev`` is (nb, nvec, 4N); returns (nb, nvec, 6N) = [e; h]."""
        t = self.torch
        ev = ev.to(device=self.device, dtype=t.complex128).contiguous()
        nvec = ev.shape[-2] if ev.dim() >= 2 else 1
        out = t.empty((self.nb, nvec, 6 * self.N), dtype=t.complex128, device=self.device)
        _check(self.lib.fame_eigen_restoration(self.h, int(nvec), ctypes.c_void_p(ev.data_ptr()),
                                               ctypes.c_void_p(out.data_ptr())))
        return out

    def _sigma_dev(self):
        if getattr(self, "_sig", None) is None:
            self._sig = self.torch.as_tensor(self.Sigma_r(), device=self.device).to(self.torch.complex128)
        return self._sig

    def _rows(self, y):
        return y.reshape(self.nb, -1)

    def curl(self, e):
        """C e for fields (nb, 3N), through the exact factorisation C = Pr diag(Sigma_r) Qr^H
        (Funhand_Generate_Lanzcos.m:218; C is DistOp.C of EvaluteError_Chiral_no_inv.m:10)."""
        return self._rows(self.Pr(self._sigma_dev() * self._rows(self.Qrs(e))))

    def curl_H(self, h):
        """C^H h = Qr diag(Sigma_r) Pr^H h (Funhand_Generate_Lanzcos.m:219)."""
        return self._rows(self.Qr(self._sigma_dev() * self._rows(self.Prs(h))))

    def EvaluteError_Chiral_no_inv(self, ev, omega, e_h):
        """(res_NFSEP, res_ORGSEP), each (nb, nvec) on the host (fun_no_inv/EvaluteError_Chiral_no_inv.m:18-36):
        res_NFSEP = ||fncB(x) - w fncA(x)||,
        res_ORGSEP = ||[N a + J b; J^H a + M b] - i w [e; h]|| with a = C^H h, b = -C e.
        ``ev`` (nb, nvec, 4N), ``omega`` (nb, nvec), ``e_h`` (nb, nvec, 6N) = [e; h] from
        Eigen_Restoration_Chiral_no_inv."""
        t = self.torch
        N3 = 3 * self.N
        ev = ev.to(device=self.device, dtype=t.complex128).reshape(self.nb, -1, 4 * self.N)
        e_h = e_h.to(device=self.device, dtype=t.complex128).reshape(self.nb, -1, 6 * self.N)
        om = np.asarray(omega, float).reshape(self.nb, -1)
        nvec = ev.shape[1]
        r_nf = np.zeros((self.nb, nvec))
        r_org = np.zeros((self.nb, nvec))
        for j in range(nvec):
            w = t.as_tensor(om[:, j], device=self.device)[:, None]
            x = ev[:, j].contiguous()
            r = self._rows(self.fncB(x)) - w * self._rows(self.fncA(x))
            r_nf[:, j] = t.linalg.vector_norm(r, dim=1).cpu().numpy()
            e = e_h[:, j, :N3].contiguous()
            h = e_h[:, j, N3:].contiguous()
            a = self.curl_H(h)
            b = -self.curl(e)
            y = self._rows(self.material_apply(t.cat([b, a], dim=1).contiguous()))  # [M b + J^H a; J b + N a]
            d1 = y[:, N3:] - 1j * w * e
            d2 = y[:, :N3] - 1j * w * h
            r_org[:, j] = t.sqrt(t.linalg.vector_norm(d1, dim=1) ** 2 + t.linalg.vector_norm(d2, dim=1) ** 2).cpu().numpy()
        return r_nf, r_org

    def set_vertex_material(self, eps, gamma, mu=1.0, lattice=None):
        """Per-vertex bi-isotropic material (eps_v, gamma_v at every vertex, x fastest) in place of
        the id table; ``lattice`` has the lattice vectors as columns (default simple cubic)."""
        e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64).reshape(-1))
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gamma, dtype=np.float64), e.shape).copy())
        if e.size != self.N:
            raise ValueError(f"expected {self.N} vertex values")
        A = np.ascontiguousarray(np.eye(3) if lattice is None else np.asarray(lattice, dtype=np.float64).reshape(3, 3))
        _check(self.lib.fame_plan_set_vertex_material(self.h, _dptr(e), _dptr(g), float(mu), _dptr(A)))

    def material_gradient(self, ev):
        """d(x^H fncB x)/d eps_v, d/d gamma_v and the vertex terms for NFSEP vectors ev
        ([nb,] nvec, 4N); returns a device tensor ([nb,] nvec, 3, N)."""
        t = self.torch
        x = ev.to(device=self.device, dtype=t.complex128).contiguous()
        squeeze = x.dim() == 2
        if squeeze:
            x = x.unsqueeze(0)
        nvec = x.shape[1]
        if x.shape[0] != self.nb or x.shape[2] != 4 * self.N:
            raise ValueError("ev must be (nb, nvec, 4N)")
        g = t.empty((self.nb, nvec, 3, self.N), dtype=t.float64, device=self.device)
        _check(self.lib.fame_material_gradient(self.h, nvec, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(g.data_ptr())))
        return g[0] if squeeze else g

    def Sigma_r(self):
        """Sigma_r = [sqrt(Lambda_q); sqrt(Lambda_q)] per item (SVD_Matrix_curl.m:74), on the host."""
        out = []
        for b in range(self.nb):
            lam = [(np.exp(1j * (2 * np.arange(n) * np.pi / n + self.theta[b, d] / n)) - 1) / self.delta[d]
                   for d, n in enumerate(self.dims)]
            n1, n2, n3 = self.dims
            lq = (np.abs(np.kron(np.ones(n3 * n2), lam[0])) ** 2
                  + np.abs(np.kron(np.ones(n3), np.kron(lam[1], np.ones(n1)))) ** 2
                  + np.abs(np.kron(lam[2], np.ones(n2 * n1))) ** 2)
            out.append(np.concatenate([np.sqrt(lq), np.sqrt(lq)]))
        return np.stack(out)


class FccPlan:
    """FCC skewed-grid operators for one grid and one k (light_locked_Huang/ FCC path).

    ``SVDSingleCurl_FCC(n1, n2, n3, vec_k, a1, a2, a3, delta)`` builds the tables
    (light_locked_Huang/SVDSingleCurl_FCC.m:1); the methods mirror
    ``p_mtx_T_prod_vec_FCC`` / ``p_mtx_TH_prod_vec_FCC`` (batched over rows),
    ``mtx_prod_vec_A_NFGEP`` (diagonal Phi), ``Compute_eigenvector``,
    ``mtx_Eps_Mu_Zeta_Xi`` + ``Construct_GEP_mtx_B`` (as operators).  Vectors are
    complex128 torch tensors on the GPU.
    """

    PARAMS = ("Eps_e", "Eps_f", "Mu_e", "Mu_f", "Zeta_e", "Zeta_f", "Xi_e", "Xi_f")

    def __init__(self, dims, vec_k, lattice, delta, device=None):
        import torch

        self.lib = load_library()
        self.torch = torch
        self.device = torch.device("cuda" if device is None else device)
        self.dims = tuple(int(d) for d in dims)
        self.n = int(np.prod(self.dims))
        arr = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))  # noqa: E731
        k, a1, a2, a3, dl = arr(vec_k), *(arr(a) for a in lattice), arr(delta)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _check(self.lib.fame_fcc_create(ctypes.byref(h), *self.dims, _dptr(k), _dptr(a1), _dptr(a2), _dptr(a3),
                                            _dptr(dl)))
        self.h = h
        _check(self.lib.fame_fcc_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)))

    def set_nfgep_variant(self, split):
        """Fused NFGEP xy kernel: both planes in LDS (0) or one plane at a time (1); same results."""
        _check(self.lib.fame_fcc_set_nfgep_variant(self.h, int(split)))

    def set_generic(self, generic=True):
        """Force the runtime-radix kernels (True) or allow the specialised ones (False)."""
        _check(self.lib.fame_fcc_set_generic(self.h, int(bool(generic))))

    def close(self):
        if getattr(self, "h", None):
            self.lib.fame_fcc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _dev(self, x, length=None):
        t = self.torch
        if not isinstance(x, t.Tensor):
            x = t.as_tensor(np.asarray(x))
        x = x.to(device=self.device, dtype=t.complex128).contiguous()
        if length is not None and x.numel() != length:
            raise ValueError(f"expected {length} complex values, got {x.numel()}")
        return x

    def _empty(self, length):
        return self.torch.empty(length, dtype=self.torch.complex128, device=self.device)

    def _vec(self, fn, x, lin, lout):
        xin = self._dev(x, lin)
        y = self._empty(lout)
        _check(fn(self.h, ctypes.c_void_p(xin.data_ptr()), ctypes.c_void_p(y.data_ptr())))
        return y

    def T(self, x):
        """p_mtx_T_prod_vec_FCC on each row of x (nvec, n) or on x (n,)."""
        return self._batched(self.lib.fame_fcc_T, x)

    def TH(self, x):
        """p_mtx_TH_prod_vec_FCC on each row of x (nvec, n) or on x (n,)."""
        return self._batched(self.lib.fame_fcc_TH, x)

    def _batched(self, fn, x):
        xin = self._dev(x)
        if xin.numel() % self.n:
            raise ValueError("length must be a multiple of n1*n2*n3")
        nvec = xin.numel() // self.n
        y = self.torch.empty_like(xin)
        _check(fn(self.h, nvec, ctypes.c_void_p(xin.data_ptr()), ctypes.c_void_p(y.data_ptr())))
        return y

    def Sigma_r(self):
        s = self.torch.empty(8 * self.n, dtype=self.torch.float64, device=self.device)
        _check(self.lib.fame_fcc_Sigma_r(self.h, ctypes.c_void_p(s.data_ptr())))
        return s

    def Pr(self, u):
        return self._vec(self.lib.fame_fcc_Pr, u, 16 * self.n, 24 * self.n)

    def Qr(self, v):
        return self._vec(self.lib.fame_fcc_Qr, v, 16 * self.n, 24 * self.n)

    def Prs(self, y):
        return self._vec(self.lib.fame_fcc_Prs, y, 24 * self.n, 16 * self.n)

    def Qrs(self, y):
        return self._vec(self.lib.fame_fcc_Qrs, y, 24 * self.n, 16 * self.n)

    def _coef(self, c):
        t = self.torch
        if np.isscalar(c) or (isinstance(c, t.Tensor) and c.numel() == 1) or np.ndim(c) == 0:
            return t.full((24 * self.n,), complex(c), dtype=t.complex128, device=self.device)
        return self._dev(c, 24 * self.n)

    def mtx_prod_vec_A_NFGEP(self, x, xi_d, zeta_d, Phi, coupling=0):
        """mtx_prod_vec_A_NFGEP (mtx_prod_vec_A_NFGEP.m:1) with diagonal Phi; xi_d, zeta_d, Phi are
        scalars or 24n vectors: the nonzeros of mtx_data.xi_d / zeta_d, on the diagonal (coupling 0)
        or in the other 12n half (coupling 1, the form ``nfgep_data`` returns)."""
        xin = self._dev(x, 32 * self.n)
        c = [self._coef(v) for v in (xi_d, zeta_d, Phi)]
        y = self._empty(32 * self.n)
        _check(self.lib.fame_fcc_nfgep_A_ex(self.h, ctypes.c_void_p(xin.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                            *(ctypes.c_void_p(v.data_ptr()) for v in c), int(coupling)))
        return y

    def Compute_eigenvector(self, x, xi_d, zeta_d, Phi, coupling=0):
        """Compute_eigenvector (Compute_eigenvector.m:1) with diagonal Phi: 32n -> 48n."""
        xin = self._dev(x, 32 * self.n)
        c = [self._coef(v) for v in (xi_d, zeta_d, Phi)]
        y = self._empty(48 * self.n)
        _check(self.lib.fame_fcc_eigenvector_ex(self.h, ctypes.c_void_p(xin.data_ptr()),
                                                ctypes.c_void_p(y.data_ptr()),
                                                *(ctypes.c_void_p(v.data_ptr()) for v in c), int(coupling)))
        return y

    def nfgep_data(self):
        """(xi_d, zeta_d, Phi) of the NFGEP for the material set by mtx_Eps_Mu_Zeta_Xi, derived from
        GEP B (Construct_GEP_mtx_B.m; mu = I, diagonal eps, zeta, xi), for coupling = 1."""
        out = [self._empty(24 * self.n) for _ in range(3)]
        _check(self.lib.fame_fcc_nfgep_data(self.h, *(ctypes.c_void_p(v.data_ptr()) for v in out)))
        return tuple(out)

    @staticmethod
    def masks(n, sets):
        """Bit masks from four index sets (X, Y, Z, vertex-or-Cubic): 0-based index arrays or
        boolean masks of length n."""
        m = np.zeros(n, np.uint8)
        for bit, s in enumerate(sets):
            s = np.asarray(s)
            sel = s.astype(bool) if (s.dtype == bool and s.size == n) else s.astype(np.int64)
            m[sel] |= np.uint8(1 << bit)
        return m

    def mtx_Eps_Mu_Zeta_Xi(self, edge_sets, face_sets, tensors_in, tensors_air):
        """Set the material of mtx_Eps_Mu_Zeta_Xi.m:1 (edge sets X, Y, Z, Cubic; face sets X, Y, Z,
        vertex; tensors (eps, mu, zeta, xi) 3x3 inside and in air)."""
        em = np.ascontiguousarray(self.masks(self.n, edge_sets))
        fm = np.ascontiguousarray(self.masks(self.n, face_sets))
        ten = np.stack([np.stack([np.asarray(T, complex).reshape(3, 3) for T in tensors_in]),
                        np.stack([np.asarray(T, complex).reshape(3, 3) for T in tensors_air])])
        tv = np.ascontiguousarray(ten.view(np.float64).reshape(-1))
        _check(self.lib.fame_fcc_set_material(self.h, em.ctypes.data_as(ctypes.c_void_p),
                                              fm.ctypes.data_as(ctypes.c_void_p), _dptr(tv)))

    def param_apply(self, name, x):
        """One of the eight matrices of mtx_Eps_Mu_Zeta_Xi.m:23-35 applied to x (12n)."""
        which = self.PARAMS.index(name)
        xin = self._dev(x, 12 * self.n)
        y = self._empty(12 * self.n)
        _check(self.lib.fame_fcc_param_apply(self.h, which, ctypes.c_void_p(xin.data_ptr()),
                                             ctypes.c_void_p(y.data_ptr())))
        return y

    def GEP_B(self, x):
        """GEP_mtx_B x (Construct_GEP_mtx_B.m:1), 48n."""
        return self._vec(self.lib.fame_fcc_gep_B, x, 48 * self.n, 48 * self.n)

"""fame_block_dot_noalloc (the allocation-free variant of fame_block_dot, its partial sums in a
file-scope device array) against fame_block_dot through the C-ABI: two calls in a row, the
second on a smaller input, bit-for-bit equal (same kernels, same order of summation); a size
beyond the static array is refused before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_block_dot_noalloc_matches_block_dot(pkg, torch_mod):
    t = torch_mod
    lib = pkg.load_library()
    S = pkg.structures
    plan = pkg.Plan((6, 6, 6), np.array([0.4, -1.3, 2.2]), S.bi_isotropic_table(13.0, 0.2, np.eye(3)),
                    S.woodpile_sc_ids((6, 6, 6)))
    rng = np.random.default_rng(17)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    for m, L in ((52, 4 * 216000), (7, 4 * 13824 + 5)):     # the second call smaller (m and len)
        print(f"block_dot_noalloc m={m} len={L}", flush=True)
        V = t.as_tensor(rng.standard_normal((m, L)) + 1j * rng.standard_normal((m, L)), device="cuda")
        w = t.as_tensor(rng.standard_normal(L) + 1j * rng.standard_normal(L), device="cuda")
        ref = t.empty(m, dtype=t.complex128, device="cuda")
        got = t.full((m,), complex(np.nan, np.nan), dtype=t.complex128, device="cuda")
        assert lib.fame_block_dot(plan.h, vp(V), m, vp(w), L, vp(ref)) == 0
        assert lib.fame_block_dot_noalloc(plan.h, vp(V), m, vp(w), L, vp(got)) == 0, lib.fame_last_error()
        plan.sync()
        assert t.equal(got, ref)
        want = (V.conj() @ w).cpu().numpy()
        assert np.abs(got.cpu().numpy() - want).max() <= 1e-12 * np.abs(want).max()
    # beyond the static array (m * ceil(len / 1024) > 262144): refused on the host
    V = t.zeros((300, 1024 * 1024), dtype=t.complex128, device="cuda")
    out = t.empty(300, dtype=t.complex128, device="cuda")
    assert lib.fame_block_dot_noalloc(plan.h, vp(V), 300, vp(V[0]), 1024 * 1024, vp(out)) == -1
    assert b"static scratch" in lib.fame_last_error()
    plan.close()

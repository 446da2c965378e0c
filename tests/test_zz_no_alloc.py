"""hplx_pdgetrs_noalloc against hplx_pdgetrs through the C ABI: the same
kernels in the same order on a device scratch array instead of the
context's buffers, so the solutions must agree bit for bit.  Two calls in a
row, the second smaller (different matrix, fewer right-hand sides), for both
A X = B and A^T X = B; the second call reuses the scratch the first one left
behind."""
import ctypes

import numpy as np
import pytest

import rochpl_amd as rp

pytestmark = pytest.mark.gpu


def _hip():
    # the HIP runtime libhplx.so is linked against (already loaded)
    rp.lib()
    return ctypes.CDLL("libamdhip64.so.7")


def _solve_both(ctx, n, nb, nrhs, shift, seed):
    rng = np.random.default_rng(seed)
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    mat = ctx.pmat_alloc(n, nb)
    out = {}
    try:
        ctx.pdrandmat(n, n, nb, mat.A, mat.ld, rp.HPL_ISEED + seed, shift)
        assert ctx.pdgetrf(rp.palg_default(), mat) == 0
        for trans in (0, 1):
            d0 = ctx.alloc(8 * n * nrhs)
            d1 = ctx.alloc(8 * n * nrhs)
            d0.upload(B)
            d1.upload(B)
            ctx.pdgetrs(mat, trans, nrhs, d0.ptr, n)
            rc = rp.lib().hplx_pdgetrs_noalloc(ctypes.byref(mat), trans, nrhs,
                                               ctypes.c_void_p(d1.ptr), n, None)
            assert rc == 0, rc
            assert _hip().hipDeviceSynchronize() == 0
            out[trans] = (d0.download((n, nrhs), np.float64),
                          d1.download((n, nrhs), np.float64))
            d0.free()
            d1.free()
    finally:
        ctx.pmat_free(mat)
    return out


@pytest.mark.parametrize("n,nb,nrhs,shift", [(3000, 512, 70, 0.0),
                                             (1100, 256, 9, 20.0)])
def test_noalloc_solve_matches_pdgetrs(ctx, n, nb, nrhs, shift):
    print("[gpu-case] pdgetrs_noalloc n=%d nb=%d nrhs=%d" % (n, nb, nrhs),
          flush=True)
    for k, (n_, nrhs_) in enumerate([(n, nrhs), (n // 2 + 1, max(nrhs // 3, 1))]):
        res = _solve_both(ctx, n_, nb, nrhs_, shift, seed=k + 1)
        for trans, (ref, got) in res.items():
            assert np.array_equal(ref.view(np.uint64), got.view(np.uint64)), \
                (n_, nrhs_, trans)


def test_noalloc_rejects_bad_arguments(ctx):
    mat = ctx.pmat_alloc(64, 32)
    try:
        L = rp.lib()
        assert L.hplx_pdgetrs_noalloc(None, 0, 1, None, 64, None) == \
            rp.HPLX_ERR_INVALID
        assert L.hplx_pdgetrs_noalloc(ctypes.byref(mat), 2, 1,
                                      ctypes.c_void_p(mat.X), 64, None) == \
            rp.HPLX_ERR_INVALID
        big = rp.PMat.from_buffer_copy(mat)
        big.nb = 2048
        assert L.hplx_pdgetrs_noalloc(ctypes.byref(big), 0, 1,
                                      ctypes.c_void_p(mat.X), 64, None) == \
            rp.HPLX_ERR_UNSUPPORTED
    finally:
        ctx.pmat_free(mat)

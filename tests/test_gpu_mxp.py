"""Mixed-precision path through the C ABI (hplx_mxp_*, hplx_spdfact).

The reference has no low-precision path.  The checks are anchored on:
  * the update: oracle/mxp_oracle.py's exact product of the rounded operands
    (pinned against torch's BF16 / e4m3 conversions), with the FP32
    accumulation error bounded;
  * the FP32 panel: the defining properties of LU with partial pivoting
    (|L| <= 1, P A = L U to FP32 rounding);
  * the solves with FP32 / FP16 factors: an FP64 CPU solve with the same
    downloaded factors;
  * the refinement residual: the FP64 oracle residual (hplo_residual);
  * end to end: the HPL scaled residual < 16 after refinement, pivots equal
    to the FP64 oracle's where the matrix fixes them (diagonal dominance).
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

import rochpl_amd as rp
from oracle import mxp_oracle as mo
from oracle import oracle as o

pytestmark = pytest.mark.gpu

U24 = 2.0 ** -24


def say(msg):
    print("[gpu-case] " + msg, flush=True)


def upload(ctx, arr):
    arr = np.asfortranarray(arr)
    d = ctx.alloc(arr.nbytes)
    d.upload(arr)
    return d


def download_factors(ctx, mat):
    dt = np.float16 if mat.storage == rp.HPLX_STORE_FP16 else np.float32
    out = np.empty((mat.ld, mat.n), dtype=dt, order="F")
    rp.lib().hplx_memcpy_d2h(ctx.h, out.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_void_p(mat.A), out.nbytes)
    return out[:mat.n, :].astype(np.float64)


def hpl_lu_solve(F, ipiv, nb, b):
    """x = U^{-1} L^{-1} P b with HPL's panel-wise (unpivoted) L, in FP64."""
    n = len(b)
    y = np.array(b, dtype=np.float64)
    for j in range(0, n, nb):
        jb = min(nb, n - j)
        for i in range(j, j + jb):
            p = ipiv[i]
            y[i], y[p] = y[p], y[i]
        y[j:j + jb] = sl.solve_triangular(F[j:j + jb, j:j + jb], y[j:j + jb],
                                          lower=True, unit_diagonal=True)
        y[j + jb:] -= F[j + jb:, j:j + jb] @ y[j:j + jb]
    return sl.solve_triangular(np.triu(F), y, lower=False)


# ---------------------------------------------------------------------------
# the low-precision trailing update
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("update,terms,storage,m,n,k", [
    (rp.HPLX_UPDATE_BF16, 1, rp.HPLX_STORE_FP32, 300, 517, 200),
    (rp.HPLX_UPDATE_BF16, 2, rp.HPLX_STORE_FP32, 300, 517, 512),
    (rp.HPLX_UPDATE_BF16, 3, rp.HPLX_STORE_FP32, 260, 130, 96),
    (rp.HPLX_UPDATE_BF16, 2, rp.HPLX_STORE_FP16, 384, 256, 256),
    (rp.HPLX_UPDATE_FP8, 1, rp.HPLX_STORE_FP32, 300, 517, 200),
    (rp.HPLX_UPDATE_FP8, 1, rp.HPLX_STORE_FP16, 513, 300, 1024),
    (rp.HPLX_UPDATE_FP32, 1, rp.HPLX_STORE_FP32, 200, 100, 64),
])
def test_mxp_update_matches_oracle(ctx, update, terms, storage, m, n, k):
    say("mxp_update upd=%d terms=%d storage=%d %dx%dx%d" % (update, terms,
                                                           storage, m, n, k))
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    ldl, ldu, ldc = m + 3, k + 5, m + 7
    L = np.zeros((ldl, k), np.float32, order="F")
    U = np.zeros((ldu, n), np.float32, order="F")
    L[:m] = rng.uniform(-1, 1, (m, k))
    # rows of U with different magnitudes exercise the per-tile scales
    U[:k] = rng.uniform(-0.5, 0.5, (k, n)) * \
        np.where(np.arange(n) % 300 < 150, 1.0, 1e-3)[None, :]
    cdt = np.float16 if storage == rp.HPLX_STORE_FP16 else np.float32
    C = np.zeros((ldc, n), cdt, order="F")
    C[:m] = rng.uniform(-1, 1, (m, n))
    dL, dU, dC = upload(ctx, L), upload(ctx, U), upload(ctx, C)
    try:
        ctx.mxp_update(update, terms, storage, m, n, k, dL.ptr, ldl, dU.ptr,
                       ldu, dC.ptr, ldc)
        out = dC.download((ldc, n), cdt)
    finally:
        for d in (dL, dU, dC):
            d.free()
    ref = C[:m].astype(np.float64) - mo.update_product(update, terms, L[:m],
                                                       U[:k])
    absp = mo.update_abs_product(update, terms, L[:m], U[:k])
    bound = U24 * (k + 8) * (np.abs(C[:m].astype(np.float64)) + absp)
    if storage == rp.HPLX_STORE_FP16:
        bound += 2.0 ** -11 * np.abs(ref)
    err = np.abs(out[:m].astype(np.float64) - ref)
    print("[mxp_update] max err / bound = %.3e" % (err / bound).max(),
          flush=True)
    assert np.all(err <= bound)
    # rows beyond m and the padding rows of C are untouched
    assert np.array_equal(out[m:], C[m:])


def test_mxp_fp8_update_exact_on_scaled_tiles(ctx):
    """Every 128 x 128 tile of L (row tile x K tile) and U (K tile x column
    tile) is an e4m3 matrix times its own power of two, all different
    (oracle scaled_e4m3_operands): the FP8 update is then exact in FP32, so
    C - L U must come back bit for bit.  A transposed, shifted or mixed-up
    exponent index in the packing or the GEMM scales some tile by a wrong
    power of two and breaks the equality."""
    L, U, sL, sU = mo.scaled_e4m3_operands()
    m, k = L.shape
    n = U.shape[1]
    say("mxp_update fp8 exact scaled tiles %dx%dx%d" % (m, n, k))
    rng = np.random.default_rng(7)
    ldl, ldu, ldc = m + 4, k + 2, m + 6
    Lp = np.zeros((ldl, k), np.float32, order="F")
    Up = np.zeros((ldu, n), np.float32, order="F")
    Lp[:m], Up[:k] = L, U
    # C: integers times 2^-(bi + 3 bj + 4) in each output tile, so that
    # C - L U is exact in FP32 as well
    C = np.zeros((ldc, n), np.float32, order="F")
    for bi in range((m + 127) // 128):
        for bj in range((n + 127) // 128):
            rs, cs = slice(bi * 128, min(m, bi * 128 + 128)), \
                slice(bj * 128, min(n, bj * 128 + 128))
            C[rs, cs] = np.ldexp(rng.integers(-1000, 1001, (rs.stop - rs.start,
                                                          cs.stop - cs.start))
                                 .astype(np.float64), -(bi + 3 * bj + 4))
    exact = C[:m].astype(np.float64) - L.astype(np.float64) @ \
        U.astype(np.float64)
    assert np.array_equal(exact.astype(np.float32).astype(np.float64), exact)
    dL, dU, dC = upload(ctx, Lp), upload(ctx, Up), upload(ctx, C)
    try:
        ctx.mxp_update(rp.HPLX_UPDATE_FP8, 1, rp.HPLX_STORE_FP32, m, n, k,
                       dL.ptr, ldl, dU.ptr, ldu, dC.ptr, ldc)
        out = dC.download((ldc, n), np.float32)
    finally:
        for d in (dL, dU, dC):
            d.free()
    bad = np.argwhere(out[:m].astype(np.float64) != exact)
    print("[mxp_update fp8 exact] %d of %d entries differ" % (len(bad),
                                                              m * n),
          flush=True)
    assert len(bad) == 0, bad[:10]
    assert np.array_equal(out[m:], C[m:])


# ---------------------------------------------------------------------------
# FP32 panel factorisation
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m,jb,nbmin", [(3000, 256, 32), (777, 96, 16),
                                        (40000, 128, 32)])
def test_spdfact_is_partial_pivoting_lu(ctx, m, jb, nbmin):
    say("spdfact m=%d jb=%d" % (m, jb))
    rng = np.random.default_rng(m + jb)
    lda = m + 13
    A = np.zeros((lda, jb), np.float32, order="F")
    A[:m] = rng.uniform(-0.5, 0.5, (m, jb))
    dA = upload(ctx, A)
    dp = ctx.alloc(4 * jb)
    try:
        info = ctx.spdfact(dA.ptr, lda, m, jb, dp.ptr, nbmin=nbmin)
        F = dA.download((lda, jb), np.float32)[:m].astype(np.float64)
        ipiv = dp.download((jb,), np.int32)
    finally:
        dA.free()
        dp.free()
    assert info == 0
    perm = np.arange(m)
    for i in range(jb):
        p = ipiv[i]
        assert i <= p < m
        perm[i], perm[p] = perm[p], perm[i]
    Lf = np.tril(F, -1)
    Lf[:jb, :jb] += np.eye(jb)
    Uf = np.triu(F[:jb])
    # partial pivoting: every multiplier is at most 1 in magnitude
    assert np.abs(np.tril(F, -1)).max() <= 1.0
    # P A = L U to FP32 rounding (backward error of Gaussian elimination)
    R = A[:m].astype(np.float64)[perm] - Lf @ Uf
    bound = 2 * (jb + 2) * U24 * (np.abs(Lf) @ np.abs(Uf))
    print("[spdfact] max |PA-LU| / bound = %.3e" % (np.abs(R) / bound).max(),
          flush=True)
    assert np.all(np.abs(R) <= bound)
    # agreement with the FP64 oracle's pivots on the same (FP32) panel:
    # columns may differ only where FP32 rounding can decide a near tie
    ref_piv, _ = o.pdfact(np.asfortranarray(A[:m].astype(np.float64)),
                          nbmin=nbmin)
    agree = np.mean(ref_piv == ipiv)
    print("[spdfact] pivot agreement with the FP64 oracle %.4f" % agree,
          flush=True)
    assert agree >= 0.9


# ---------------------------------------------------------------------------
# solves with the stored low-precision factors, refinement residual
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("solve_fp64", [1, 0])
@pytest.mark.parametrize("storage,update,n,nb,shift", [
    (rp.HPLX_STORE_FP32, rp.HPLX_UPDATE_FP32, 2048, 256, 0.0),
    (rp.HPLX_STORE_FP16, rp.HPLX_UPDATE_FP8, 1536, 512, 384.0),
    # a last block smaller than NB (its inverse and interchanges apart)
    (rp.HPLX_STORE_FP32, rp.HPLX_UPDATE_FP32, 1000, 128, 0.0),
    (rp.HPLX_STORE_FP16, rp.HPLX_UPDATE_FP8, 1300, 256, 325.0),
])
def test_mxp_getrs_and_residual(ctx, storage, update, n, nb, shift,
                                solve_fp64):
    """The solve with the stored factors: in FP64 it equals the same solve on
    the CPU to rounding; in FP32 (the default) it is backward stable, its x
    solving L U x = P b to within a multiple of n * eps32 relative to
    |L| |U| |x|, a bound an indexing or interchange mistake exceeds by
    orders of magnitude."""
    say("mxp getrs storage=%d n=%d nb=%d solve_fp64=%d"
        % (storage, n, nb, solve_fp64))
    mat = ctx.mxp_alloc(n, nb, storage)
    opts = rp.mxp_opts_default(update, 1)
    opts.solve_fp64 = solve_fp64
    try:
        ctx.mxp_matgen(mat, rp.HPL_ISEED, shift)
        # the stored matrix is the generator's, rounded to the storage type
        A64 = o.matgen(n, n, diag_shift=shift)
        A0 = download_factors(ctx, mat)
        if storage == rp.HPLX_STORE_FP32:
            assert np.array_equal(A0, A64.astype(np.float32).astype(np.float64))
        else:
            want = A64.astype(np.float32).astype(np.float16).astype(np.float64)
            bad = np.argwhere(A0 != want)
            if len(bad):
                one = A64.astype(np.float16).astype(np.float64)
                sub = np.abs(want[tuple(bad.T)]) < 2.0 ** -14
                print("[mxp] fp16 matgen: %d mismatches, %d subnormal, %d equal "
                      "to a single double->half rounding; first %s: gpu %r "
                      "want %r fp64 %r" % (
                          len(bad), int(sub.sum()),
                          int(np.sum(A0[tuple(bad.T)] == one[tuple(bad.T)])),
                          bad[0], A0[tuple(bad[0])], want[tuple(bad[0])],
                          A64[tuple(bad[0])]), flush=True)
            assert len(bad) == 0
        assert ctx.mxp_pdgetrf(rp.palg_default(), opts, mat) == 0
        F = download_factors(ctx, mat)
        ipiv = ctx.download_vec(mat.ipiv, n, np.int32)
        b = o.matgen_tile(n, 0, n, n, 1)[:, 0]
        d = upload(ctx, b)
        ctx.mxp_pdgetrs(mat, d.ptr)
        x = d.download((n,), np.float64)
        if solve_fp64:
            # the same solve on the CPU in FP64 with the downloaded factors
            xr = hpl_lu_solve(F, ipiv, nb, b)
            assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-12
        else:
            # P b with HPL's panel-wise interchanges, then the backward error
            # of the FP32 solve, all in FP64 on the CPU
            pb = np.array(b, dtype=np.float64)
            for i in range(n):
                pb[i], pb[ipiv[i]] = pb[ipiv[i]], pb[i]
            L = np.tril(F, -1) + np.eye(n)
            U = np.triu(F)
            res_b = np.abs(pb - L @ (U @ x))
            scale = (np.abs(L) @ (np.abs(U) @ np.abs(x))) + np.abs(pb)
            rel = np.max(res_b / scale)
            eps32 = np.finfo(np.float32).eps
            print("[mxp] fp32 solve backward error %.3e (n eps32 %.3e)"
                  % (rel, n * eps32), flush=True)
            assert rel <= n * eps32
            x32 = x.astype(np.float32).astype(np.float64)
            assert np.array_equal(x, x32)  # an FP32 solution, widened
        # refinement residual against the FP64 oracle's
        res = ctx.mxp_residual(mat, d.ptr)
        ref = o.residual(n, x, diag_shift=shift)
        d.free()
        # both sum b - A x in FP64 in different orders: they may differ by
        # about eps * ||A||_oo ||x||_oo, whatever the size of r itself
        scale = np.finfo(float).eps * ref["AnormI"] * ref["XnormI"]
        assert abs(res["rnorm"] - ref["rnorm"]) <= 4 * scale
        assert res["AnormI"] == pytest.approx(ref["AnormI"], rel=1e-13)
        assert res["BnormI"] == ref["BnormI"]
        assert res["XnormI"] == ref["XnormI"]
    finally:
        ctx.mxp_free(mat)


# ---------------------------------------------------------------------------
# end to end: factor + FP64 refinement
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("storage,update,terms,n,nb,shift", [
    (rp.HPLX_STORE_FP32, rp.HPLX_UPDATE_FP32, 1, 4096, 256, 0.0),
    (rp.HPLX_STORE_FP32, rp.HPLX_UPDATE_BF16, 2, 4096, 256, 0.0),
    (rp.HPLX_STORE_FP32, rp.HPLX_UPDATE_BF16, 3, 4096, 512, 0.0),
    (rp.HPLX_STORE_FP32, rp.HPLX_UPDATE_BF16, 1, 3000, 256, 750.0),
    (rp.HPLX_STORE_FP16, rp.HPLX_UPDATE_FP8, 1, 4096, 1024, 1024.0),
    (rp.HPLX_STORE_FP16, rp.HPLX_UPDATE_BF16, 2, 2500, 256, 625.0),
])
@pytest.mark.parametrize("solve_fp64", [0, 1])
def test_mxp_pdgesv_converges(ctx, storage, update, terms, n, nb, shift,
                              solve_fp64):
    say("mxp_pdgesv storage=%d update=%d terms=%d n=%d nb=%d shift=%g "
        "solve_fp64=%d" % (storage, update, terms, n, nb, shift, solve_fp64))
    mat = ctx.mxp_alloc(n, nb, storage)
    opts = rp.mxp_opts_default(update, terms)
    opts.solve_fp64 = solve_fp64
    try:
        ctx.mxp_matgen(mat, rp.HPL_ISEED, shift)
        rc, res = ctx.mxp_pdgesv(rp.palg_default(), opts, mat)
        x = ctx.download_vec(mat.X, n)
        ipiv = ctx.download_vec(mat.ipiv, n, np.int32)
    finally:
        ctx.mxp_free(mat)
    print("[mxp] its=%d resid=%.4f history=%s" % (
        res["its"], res["resid"], ["%.3g" % h for h in res["history"]]),
        flush=True)
    assert rc == 0 and res["converged"] == 1
    ref = o.residual(n, x, diag_shift=shift)
    assert ref["resid"] < 16.0
    # the device's and the oracle's ||b - A x||_oo differ by rounding of
    # order eps ||A|| ||x||, i.e. well below one unit of the scaled residual
    assert abs(ref["resid"] - res["resid"]) <= 0.1
    r64 = o.pdgesv(n, nb, diag_shift=shift)
    err = np.linalg.norm(x - r64["x"], np.inf) / np.linalg.norm(r64["x"], np.inf)
    # forward error <= cond * backward error; the scaled residual s bounds
    # the backward error by s * n * eps (for both x and the FP64 solution)
    kappa = np.linalg.cond(o.matgen(n, n, diag_shift=shift), np.inf)
    bound = kappa * (max(res["resid"], 0.1) + 0.1) * n * np.finfo(float).eps
    print("[mxp] ||x - x_fp64|| / ||x_fp64|| = %.3e (cond_oo %.3g, bound %.3e)"
          % (err, kappa, bound), flush=True)
    assert err <= bound
    if shift >= n / 4:
        # diagonally dominant: partial pivoting keeps every diagonal pivot,
        # as the FP64 oracle does
        assert np.array_equal(ipiv, np.arange(n))
        assert np.array_equal(ipiv, r64["ipiv"])


@pytest.mark.parametrize("storage,update,terms,n,nb,shift", [
    (rp.HPLX_STORE_FP16, rp.HPLX_UPDATE_FP8, 1, 8192, 1024, 2048.0),
    (rp.HPLX_STORE_FP32, rp.HPLX_UPDATE_BF16, 3, 6000, 512, 0.0),
    (rp.HPLX_STORE_FP32, rp.HPLX_UPDATE_BF16, 1, 5000, 256, 1250.0),
    (rp.HPLX_STORE_FP16, rp.HPLX_UPDATE_BF16, 2, 5000, 256, 1250.0),
])
def test_mxp_lookahead_is_bit_identical_to_serial(ctx, storage, update, terms,
                                                  n, nb, shift):
    """Depth-1 look-ahead of the mixed-precision loop (panel k+1 factored on
    the panel stream beside the claimed / resumed rest of update k) against
    every step in order on one stream: the same pivots, the same stored
    factors, the same x and the same refinement history, bit for bit.  6000
    and 5000 leave a ragged last panel."""
    say("mxp look-ahead vs serial storage=%d update=%d terms=%d n=%d nb=%d"
        % (storage, update, terms, n, nb))
    res = {}
    mat = ctx.mxp_alloc(n, nb, storage)
    opts = rp.mxp_opts_default(update, terms)
    try:
        for la in (False, True):
            ctx.lookahead = la
            ctx.mxp_matgen(mat, rp.HPL_ISEED, shift)
            rc, r = ctx.mxp_pdgesv(rp.palg_default(), opts, mat)
            assert rc == 0 and r["converged"] == 1
            res[la] = (ctx.download_vec(mat.ipiv, n, np.int32),
                       ctx.download_vec(mat.X, n), download_factors(ctx, mat),
                       r)
    finally:
        ctx.lookahead = True
        ctx.mxp_free(mat)
    (p0, x0, f0, r0), (p1, x1, f1, r1) = res[False], res[True]
    assert np.array_equal(p0, p1)
    assert np.array_equal(f0.view(np.uint64), f1.view(np.uint64))
    assert np.array_equal(x0.view(np.uint64), x1.view(np.uint64))
    assert r0["its"] == r1["its"] and r0["history"] == r1["history"]
    print("[mxp lookahead] identical; its %d resid %.4f; factor %.1f ms serial, "
          "%.1f ms look-ahead" % (r1["its"], r1["resid"], r0["factor_ms"],
                                  r1["factor_ms"]), flush=True)


@pytest.mark.parametrize("pf,rf", [(rp.HPLX_CROUT, rp.HPLX_LEFT_LOOKING),
                                   (rp.HPLX_LEFT_LOOKING, rp.HPLX_CROUT)])
@pytest.mark.parametrize("storage,update,terms,n,nb,shift", [
    (rp.HPLX_STORE_FP32, rp.HPLX_UPDATE_BF16, 3, 4096, 512, 0.0),
    (rp.HPLX_STORE_FP16, rp.HPLX_UPDATE_FP8, 1, 4096, 1024, 1024.0),
])
def test_mxp_panel_variants(ctx, pf, rf, storage, update, terms, n, nb, shift):
    """The FP32 panel in the left-looking / Crout orders (the FP64 path's
    variant loop): the refinement converges to the HPL criterion, and the
    pivots agree with the right-looking run's except where FP32 rounding
    decides a near tie (none on the diagonally dominant matrix)."""
    say("mxp variants pfact=%d rfact=%d storage=%d update=%d n=%d"
        % (pf, rf, storage, update, n))
    mat = ctx.mxp_alloc(n, nb, storage)
    opts = rp.mxp_opts_default(update, terms)
    piv = {}
    try:
        for variant in ((rp.HPLX_RIGHT_LOOKING, rp.HPLX_RIGHT_LOOKING), (pf, rf)):
            palg = rp.palg_default()
            palg.pfact, palg.rfact = variant
            ctx.mxp_matgen(mat, rp.HPL_ISEED, shift)
            rc, res = ctx.mxp_pdgesv(palg, opts, mat)
            assert rc == 0 and res["converged"] == 1
            x = ctx.download_vec(mat.X, n)
            assert o.residual(n, x, diag_shift=shift)["resid"] < 16.0
            piv[variant] = ctx.download_vec(mat.ipiv, n, np.int32)
    finally:
        ctx.mxp_free(mat)
    p0, p1 = piv.values()
    agree = float(np.mean(p0 == p1))
    print("[mxp variants] pivot agreement with right-looking %.4f" % agree,
          flush=True)
    if shift >= n / 4:
        assert np.array_equal(p1, np.arange(n))
    else:
        assert agree >= 0.99


def test_mxp_rejects_bad_options(ctx):
    mat = ctx.mxp_alloc(256, 64, rp.HPLX_STORE_FP16)
    try:
        with pytest.raises(rp.HplxError):
            ctx.mxp_pdgetrf(rp.palg_default(),
                            rp.mxp_opts_default(rp.HPLX_UPDATE_FP32, 1), mat)
        with pytest.raises(rp.HplxError):
            ctx.mxp_pdgetrf(rp.palg_default(),
                            rp.mxp_opts_default(rp.HPLX_UPDATE_BF16, 4), mat)
        bad = rp.mxp_opts_default(rp.HPLX_UPDATE_BF16, 2)
        bad.solve_fp64 = 2
        with pytest.raises(rp.HplxError):
            ctx.mxp_pdgetrf(rp.palg_default(), bad, mat)
    finally:
        ctx.mxp_free(mat)

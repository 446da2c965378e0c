"""The persistent panel kernel (HPLX_OPT_PANEL_KERNEL = 1): each FP64 panel
factorisation as one launch that walks the recursion of HPL_pdrpan*T /
HPL_pdpan*T (src/pfact/HPL_pdrpanrlT.cpp:19-161) with grid barriers between
its steps.  Checked against the oracle's restatement of the same variant
(pivots bit-exact, factors to rounding), against the launch sequence it
replaces, and inside the whole solve with and without the look-ahead."""
import numpy as np
import pytest

import rochpl_amd as rp
from oracle import oracle as o

pytestmark = pytest.mark.gpu

_FACT = {"L": rp.HPLX_LEFT_LOOKING, "C": rp.HPLX_CROUT,
         "R": rp.HPLX_RIGHT_LOOKING}


def say(msg):
    print("[gpu-case] " + msg, flush=True)


@pytest.fixture
def pk(ctx):
    ctx.panel_kernel = 1  # every panel
    yield ctx
    ctx.panel_kernel = 2  # the default


def _even_ld(m):
    # the streamed product tile wants an even leading dimension
    return m + 10 + (m & 1)


def _pdfact(ctx, P0, m, jb, a):
    lda = _even_ld(m)
    buf = np.zeros((lda, jb), dtype=np.float64, order="F")
    buf[:m] = P0
    dP = ctx.alloc(buf.nbytes)
    dP.upload(buf)
    dpiv = ctx.alloc(4 * jb)
    info = ctx.pdfact_palg(a, dP.ptr, lda, m, jb, dpiv.ptr, 0)
    return (info, dP.download((lda, jb), np.float64)[:m],
            dpiv.download(jb, np.int32))


# 3001: odd panel height (the products' last row pair); 70000: one leaf over
# 46 workgroups; 900 x 128 with nbmin 64: leaves blocked in 32-column
# chunks; 5 x 3: a single workgroup
@pytest.mark.parametrize("m,jb,nbmin,ndiv", [
    (3000, 256, 32, 2), (3001, 256, 32, 2), (20000, 64, 32, 2),
    (1500, 96, 16, 3), (900, 128, 64, 2), (5, 3, 32, 2), (70000, 32, 32, 2)])
def test_panel_kernel_pdfact_matches_oracle(pk, m, jb, nbmin, ndiv):
    say("panel kernel pdfact m=%d jb=%d nbmin=%d ndiv=%d" % (m, jb, nbmin, ndiv))
    P0 = np.asfortranarray(o.matgen_tile(m + 11, 0, 0, m, jb, seed=m + jb))
    a = rp.palg_default()
    a.nbmin, a.nbdiv = nbmin, ndiv
    info, got, gpiv = _pdfact(pk, P0, m, jb, a)
    assert pk.panel_kernel_runs() == 1
    P = P0.copy(order="F")
    wpiv, winfo = o.pdfact(P, nbmin=nbmin, ndiv=ndiv)
    assert info == winfo == 0
    assert np.array_equal(gpiv, wpiv)
    assert np.max(np.abs(got - P)) < 1e-11


@pytest.mark.parametrize("pf", ["L", "C", "R"])
@pytest.mark.parametrize("rf", ["L", "C", "R"])
@pytest.mark.parametrize("m,jb,nbmin,ndiv", [
    (3000, 256, 16, 2), (2000, 192, 48, 3)])
def test_panel_kernel_variants_match_oracle(pk, pf, rf, m, jb, nbmin, ndiv):
    say("panel kernel variant pfact=%s rfact=%s m=%d jb=%d nbmin=%d ndiv=%d"
        % (pf, rf, m, jb, nbmin, ndiv))
    P0 = np.asfortranarray(o.matgen_tile(m + 11, 0, 0, m, jb, seed=m + jb))
    a = rp.palg_default()
    a.pfact, a.rfact, a.nbmin, a.nbdiv = _FACT[pf], _FACT[rf], nbmin, ndiv
    info, got, gpiv = _pdfact(pk, P0, m, jb, a)
    assert pk.panel_kernel_runs() == 1
    P = P0.copy(order="F")
    wpiv, winfo = o.pdfact(P, nbmin=nbmin, ndiv=ndiv, pfact=_FACT[pf],
                           rfact=_FACT[rf])
    assert info == winfo == 0
    assert np.array_equal(gpiv, wpiv)
    assert np.max(np.abs(got - P)) < 1e-11


@pytest.mark.parametrize("m,jb", [(12000, 512), (3001, 896)])
def test_panel_kernel_against_launch_sequence(ctx, m, jb):
    """The same panel factored both ways: equal pivots and factors, bit for
    bit (products under 128 columns take the generic GEMM in the launch
    sequence and the streamed tile in the kernel)."""
    say("panel kernel vs launches m=%d jb=%d" % (m, jb))
    P0 = np.asfortranarray(o.matgen_tile(m + 11, 0, 0, m, jb, seed=7 * m + jb))
    out = {}
    try:
        for on in (False, True):
            ctx.panel_kernel = on
            out[on] = _pdfact(ctx, P0, m, jb, rp.palg_default())
            assert ctx.panel_kernel_runs() == (1 if on else 0)
    finally:
        ctx.panel_kernel = 2
    (i0, f0, p0), (i1, f1, p1) = out[False], out[True]
    assert i0 == i1 == 0
    assert np.array_equal(p0, p1)
    # the streamed tile and the generic GEMM form every product in the same
    # MFMA order, so the factors are the same bits
    assert np.array_equal(f0.view(np.uint64), f1.view(np.uint64))


@pytest.mark.parametrize("n,nb", [(4096, 256), (2000, 896), (1000, 64)])
def test_panel_kernel_pdgesv_matches_oracle(pk, n, nb):
    say("panel kernel pdgesv n=%d nb=%d" % (n, nb))
    mat = pk.pmat_alloc(n, nb)
    try:
        pk.pdrandmat(n, n + 1, nb, mat.A, mat.ld, rp.HPL_ISEED, 0.0)
        assert pk.pdgesv(rp.palg_default(), mat) == 0
        runs = pk.panel_kernel_runs()
        x = pk.download_vec(mat.X, n)
        ipiv = pk.download_vec(mat.ipiv, n, np.int32)
        chk = pk.pdtest_check(mat)
    finally:
        pk.pmat_free(mat)
    assert runs == (n + nb - 1) // nb
    ref = o.pdgesv(n, nb)
    assert np.array_equal(ipiv, ref["ipiv"])
    assert np.max(np.abs(x - ref["x"])) < 1e-8 * np.abs(ref["x"]).max()
    assert chk["resid"] < 16.0


@pytest.mark.parametrize("n,nb", [(16384, 512), (9001, 256)])
def test_panel_kernel_lookahead_is_bit_identical_to_serial(pk, n, nb):
    """The persistent panel beside the claimed update (look-ahead) against
    every step in order: identical pivots, factors and x; the pivots also
    equal those of the launch sequence.  9001: a ragged last panel of 41
    columns."""
    say("panel kernel look-ahead vs serial n=%d nb=%d" % (n, nb))
    res = {}
    mat = pk.pmat_alloc(n, nb)
    try:
        for on, la in ((True, False), (True, True), (False, True)):
            pk.panel_kernel = on
            pk.lookahead = la
            pk.pdrandmat(n, n + 1, nb, mat.A, mat.ld, rp.HPL_ISEED, 0.0)
            assert pk.pdgesv(rp.palg_default(), mat) == 0
            res[(on, la)] = (pk.download_vec(mat.ipiv, n, np.int32),
                             pk.download_vec(mat.X, n),
                             pk.download_matrix(mat, rows=n, cols=n + 1),
                             pk.pdtest_check(mat)["resid"],
                             pk.panel_kernel_runs())
    finally:
        pk.lookahead = True
        pk.pmat_free(mat)
    (p0, x0, f0, r0, k0) = res[(True, False)]
    (p1, x1, f1, r1, k1) = res[(True, True)]
    (p2, _, _, r2, k2) = res[(False, True)]
    assert k0 == k1 and k0 >= n // nb and k2 == 0
    assert np.array_equal(p0, p1) and np.array_equal(p0, p2)
    assert np.array_equal(x0.view(np.uint64), x1.view(np.uint64))
    assert np.array_equal(np.asarray(f0).view(np.uint64),
                          np.asarray(f1).view(np.uint64))
    assert r0 < 16.0 and r1 < 16.0 and r2 < 16.0
    print("[panel-kernel] n=%d nb=%d look-ahead identical; resid %.4f"
          % (n, nb, r1), flush=True)


def test_panel_kernel_default_mixes_forms_bit_identically(ctx):
    """The default (2) at N = 66560, NB = 512: the first panels (66560,
    66048 and 65536 rows) run as the kernel, the rest as the launch sequence.  Against
    the launch sequence throughout, and against the serial schedule: the same
    pivots and the same x, bit for bit; the residual passes."""
    n, nb = 66560, 512
    say("panel kernel default, mixed forms n=%d nb=%d" % (n, nb))
    res = {}
    mat = ctx.pmat_alloc(n, nb)
    try:
        for mode, la in ((2, True), (0, True), (2, False)):
            ctx.panel_kernel = mode
            ctx.lookahead = la
            ctx.pdrandmat(n, n + 1, nb, mat.A, mat.ld, rp.HPL_ISEED, 0.0)
            assert ctx.pdgesv(rp.palg_default(), mat) == 0
            res[(mode, la)] = (ctx.download_vec(mat.ipiv, n, np.int32),
                               ctx.download_vec(mat.X, n),
                               ctx.pdtest_check(mat)["resid"],
                               ctx.panel_kernel_runs())
    finally:
        ctx.panel_kernel = 2
        ctx.lookahead = True
        ctx.pmat_free(mat)
    (p0, x0, r0, k0) = res[(2, True)]
    (p1, x1, r1, k1) = res[(0, True)]
    (p2, x2, r2, k2) = res[(2, False)]
    tall = sum(1 for k in range((n + nb - 1) // nb) if n - k * nb >= 65536)
    assert k0 == k2 == tall == 3 and k1 == 0
    assert np.array_equal(p0, p1) and np.array_equal(p0, p2)
    assert np.array_equal(x0.view(np.uint64), x1.view(np.uint64))
    assert np.array_equal(x0.view(np.uint64), x2.view(np.uint64))
    assert max(r0, r1, r2) < 16.0


@pytest.mark.parametrize("case", ["many_steps", "odd_ld", "odd_width"])
def test_panel_kernel_falls_back_to_the_launch_sequence(pk, case):
    """Panels the kernel does not take run as the launch sequence, with the
    oracle's pivots: more than 512 recursion steps (NBMIN 1 on 600
    columns), an odd leading dimension (the streamed product's pairs), a
    product depth that is not a multiple of 8 (41 columns at NBMIN 20: a
    product of depth 20)."""
    m, jb, nbmin = {"many_steps": (2000, 600, 1), "odd_ld": (3000, 256, 32),
                    "odd_width": (3000, 41, 20)}[case]
    say("panel kernel fallback %s m=%d jb=%d nbmin=%d" % (case, m, jb, nbmin))
    a = rp.palg_default()
    a.nbmin = nbmin
    steps = rp.panel_steps(a, jb)
    if case == "many_steps":
        assert len(steps) > 512
    if case == "odd_width":
        assert any(st[0] == rp.PANEL_GEMM and st[4] % 8 for st in steps)
    P0 = np.asfortranarray(o.matgen_tile(m + 11, 0, 0, m, jb, seed=m + jb))
    lda = m + 11 if case == "odd_ld" else _even_ld(m)
    assert (lda % 2 == 1) == (case == "odd_ld")
    buf = np.zeros((lda, jb), dtype=np.float64, order="F")
    buf[:m] = P0
    dP = pk.alloc(buf.nbytes)
    dP.upload(buf)
    dpiv = pk.alloc(4 * jb)
    info = pk.pdfact_palg(a, dP.ptr, lda, m, jb, dpiv.ptr, 0)
    assert pk.panel_kernel_runs() == 0
    got = dP.download((lda, jb), np.float64)[:m]
    P = P0.copy(order="F")
    wpiv, winfo = o.pdfact(P, nbmin=nbmin, ndiv=2)
    assert info == winfo == 0
    assert np.array_equal(dpiv.download(jb, np.int32), wpiv)
    assert np.max(np.abs(got - P)) < 1e-11


def test_panel_kernel_option_roundtrip(ctx):
    assert ctx.panel_kernel == 2  # the default: panels of >= 65536 rows
    for v in (0, 1, 2):
        ctx.panel_kernel = v
        assert ctx.panel_kernel == v
    for bad in (-1, 3):
        assert rp.lib().hplx_set_option(ctx.h, rp.HPLX_OPT_PANEL_KERNEL,
                                        bad) == rp.HPLX_ERR_INVALID
    assert ctx.panel_kernel == 2


def test_panel_kernel_default_takes_tall_panels_only(ctx):
    """The default (2): a panel of 65536 rows or more runs as the kernel,
    a shorter one as the launch sequence; both give the same bits."""
    out = {}
    for m in (65535, 65536):
        say("panel kernel default m=%d" % m)
        P0 = np.asfortranarray(o.matgen_tile(m + 11, 0, 0, m, 64, seed=m))
        out[m] = _pdfact(ctx, P0, m, 64, rp.palg_default())
        assert out[m][0] == 0
        assert ctx.panel_kernel_runs() == (1 if m >= 65536 else 0)

"""Pins the CPU oracle (oracle/hpl_oracle.c) against checks that do not
restate its own formulas: LAPACK dgetrf pivots and solves (scipy), exact
rational arithmetic for the generator's rounding, sequential LCG stepping for
the O(log k) jump, brute-force interchanges for the pivot composition, and
numpy norms for the residual check.  CPU only."""
import fractions
import random

import numpy as np
import pytest
import scipy.linalg as sl
from scipy.linalg import lapack

from oracle import oracle as o

MULT = 6364136223846793005
M64 = (1 << 64) - 1


def seq_state(seed, k):
    s = (MULT * seed + 1) & M64
    for _ in range(k):
        s = (MULT * s + 1) & M64
    return s


def exact_value(s):
    # RN(0.5 - RN(s * 2^-32) * 2^-32): float(Fraction) rounds correctly once,
    # the power-of-two product is exact and the subtraction rounds once.
    y = float(fractions.Fraction(s, 1 << 32))
    return 0.5 - y * 2.0 ** -32


def test_matgen_spot_values_seed100():
    # a(0..3, 0) for seed 100 as derived in the survey of the reference
    want = [-5.159944193766641e-05, 0.38413122531282284, 0.4436836675989427,
            -0.46046395121548966]
    A = o.matgen(4, 1)
    assert A[:, 0].tolist() == want


def test_jump_equals_sequential_stepping():
    rng = random.Random(1)
    for k in [0, 1, 2, 3, 7, 64, 1000, 4097] + [rng.randrange(20000)
                                               for _ in range(10)]:
        assert o.state(100, k) == seq_state(100, k)
    assert o.state(7, 12345) == seq_state(7, 12345)


def test_value_formula_matches_exact_rounding():
    rng = random.Random(2)
    states = [0, 1, M64, 1 << 63, (1 << 63) - 1, 0xFFFFFFFF, 1 << 32]
    states += [rng.getrandbits(64) for _ in range(20000)]
    for s in states:
        assert o.value(s) == exact_value(s), hex(s)


def test_matgen_layout_is_state_i_plus_j_times_m():
    m, n = 37, 5
    A = o.matgen(m, n, seed=42)
    for (i, j) in [(0, 0), (36, 0), (0, 4), (17, 3), (36, 4)]:
        assert A[i, j] == exact_value(seq_state(42, i + j * m))
    # tile generator agrees with the whole-matrix one
    T = o.matgen_tile(m, 5, 2, 20, 3, seed=42)
    assert np.array_equal(T, A[5:25, 2:5])


def test_matgen_diag_shift():
    A = o.matgen(6, 7)
    B = o.matgen(6, 7, diag_shift=3.0)
    D = B - A
    assert np.allclose(D[:, :6], 3.0 * np.eye(6)) and np.all(D[:, 6] == 0)


@pytest.mark.parametrize("n,nb", [(8, 4), (64, 32), (257, 64), (1000, 64),
                                  (300, 100), (130, 7)])
def test_pdgesv_pivots_equal_lapack(n, nb):
    r = o.pdgesv(n, nb)
    A = o.matgen(n, n + 1)
    lu, piv = sl.lu_factor(A[:, :n])
    assert r["info"] == 0
    assert np.array_equal(r["ipiv"], piv)
    x = sl.lu_solve((lu, piv), A[:, n])
    assert np.max(np.abs(r["x"] - x)) <= 1e-9 * np.max(np.abs(x))
    # U equals LAPACK's U to rounding (same pivots => same factors)
    U = np.triu(r["LU"][:, :n])
    assert np.allclose(U, np.triu(lu), rtol=0, atol=1e-10 * np.abs(lu).max())


@pytest.mark.parametrize("nbmin,ndiv", [(1, 2), (4, 3), (32, 2), (64, 2)])
def test_pdgesv_recursion_params_keep_pivots(nbmin, ndiv):
    n, nb = 200, 48
    r = o.pdgesv(n, nb, nbmin=nbmin, ndiv=ndiv)
    _, piv = sl.lu_factor(o.matgen(n, n + 1)[:, :n])
    assert np.array_equal(r["ipiv"], piv)


def test_pdfact_tall_panel_matches_dgetrf():
    m, jb = 700, 96
    P = np.asfortranarray(o.matgen_tile(1000, 0, 0, m, jb, seed=5))
    ref, piv, info = lapack.dgetrf(P.copy())
    ipiv, inf = o.pdfact(P, nbmin=32, ndiv=2)
    assert inf == 0 and info == 0
    assert np.array_equal(ipiv, piv)
    assert np.allclose(P, ref, rtol=0, atol=1e-12)


def brute_apply(A, ipiv):
    A = A.copy()
    for i, p in enumerate(ipiv):
        A[[i, p], :] = A[[p, i], :]
    return A


@pytest.mark.parametrize("jb,m,seed", [(1, 5, 0), (8, 8, 1), (16, 100, 2),
                                       (64, 300, 3), (128, 5000, 4)])
def test_pipid_equals_bruteforce(jb, m, seed):
    rng = np.random.default_rng(seed)
    ipiv = np.array([rng.integers(i, m) for i in range(jb)], dtype=np.int32)
    # force repeated outside targets and inside targets
    if jb > 4:
        ipiv[1] = ipiv[0] if ipiv[0] >= 1 else 1
        ipiv[2] = 3
    permU = o.pipid(ipiv)
    A = np.asfortranarray(rng.standard_normal((m, 7)))
    want = brute_apply(A, ipiv)
    B = A.copy(order="F")
    o.dlaswp00N(B, permU)
    assert np.array_equal(B, want)
    C = A.copy(order="F")
    o.laswp(C, ipiv)
    assert np.array_equal(C, want)
    # format: top sources, displaced destinations or -1
    assert np.all(permU[jb:] >= -1)


def test_residual_matches_numpy():
    n = 300
    r = o.pdgesv(n, 32)
    A = o.matgen(n, n + 1)
    x = r["x"]
    res = o.residual(n, x)
    rr = A[:, n] - A[:, :n] @ x
    an_i = np.abs(A[:, :n]).sum(axis=1).max()
    an_1 = np.abs(A[:, :n]).sum(axis=0).max()
    eps = 2.0 ** -53
    assert res["AnormI"] == pytest.approx(an_i, rel=1e-13)
    assert res["Anorm1"] == pytest.approx(an_1, rel=1e-13)
    assert res["BnormI"] == np.abs(A[:, n]).max()
    assert res["XnormI"] == np.abs(x).max()
    # r is rounding noise; two summation orders agree to a fraction of the
    # scaled-residual unit eps*(||A||oo*||x||oo+||b||oo)*n
    unit = eps * (an_i * np.abs(x).max() + np.abs(A[:, n]).max()) * n
    assert abs(res["rnorm"] - np.abs(rr).max()) < 0.5 * unit
    assert abs(res["r0"] - abs(rr[0])) < 0.5 * unit
    want = np.abs(rr).max() / unit
    assert abs(res["resid"] - want) < 0.5
    assert res["resid"] == pytest.approx(res["rnorm"] / unit, rel=1e-12)
    assert res["resid"] < 16.0
    # the |r_0| form never exceeds the true one
    assert res["resid_r0_form"] <= res["resid"]


def test_residual_flags_a_wrong_solution():
    n = 100
    x = o.pdgesv(n, 16)["x"].copy()
    x[3] += 1e-6
    assert o.residual(n, x)["resid"] > 16.0


def test_zero_pivot_reported_lapack_style():
    n = 40
    A = o.matgen(n, n + 1).copy()
    A[:, 5] = 0.0  # exact zero column -> zero pivot at column 6 (1-based)
    r = o.pdgesv(n, 8, Ab=A)
    _, _, info = lapack.dgetrf(A[:, :n].copy())
    assert r["info"] == 6 == info


def test_flop_counts():
    n = 4096
    assert o.flops_ref(n) == pytest.approx(2 / 3 * n ** 3 + 1.5 * n ** 2)
    assert o.flops_hpl(n) == pytest.approx(2 / 3 * n ** 3 + 2 * n ** 2)


def test_hpl_ld_rule():
    assert o.hpl_ld(193280) == 193312
    assert o.hpl_ld(4096) == 4128

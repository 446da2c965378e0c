"""Pins the oracle's panel variants (left-looking, Crout, right-looking for
both the leaf and the recursion; src/HPL_pddriver.cpp:164-200) on the CPU.

Two checks fixed by the mathematics, not by the oracle's formulas:
  * with its sequential loops (no host BLAS) every element receives the
    products l_ik * u_kj in the order k = 0, 1, ... under every variant --
    only the moment of each update differs -- so all nine (pfact, rfact)
    pairs must give the same factors and pivots bit for bit, at every
    blocking.  A variant that drops, repeats or misplaces an update (a
    wrong chunk offset, a missing solve) breaks the equality;
  * the pivots equal LAPACK dgetrf's (partial pivoting, first row on ties)
    and the factors its factors to rounding.
"""
import itertools

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as o

FACTS = (o.LEFT_LOOKING, o.CROUT, o.RIGHT_LOOKING)


def _panel(m, jb, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.uniform(-0.5, 0.5, (m, jb)))


@pytest.mark.parametrize("m,jb,nbmin,ndiv", [
    (300, 96, 8, 2), (257, 64, 16, 3), (200, 50, 4, 2), (120, 120, 32, 2),
    (90, 33, 1, 2), (40, 40, 64, 2)])
def test_all_variants_agree_bit_for_bit(m, jb, nbmin, ndiv):
    P0 = _panel(m, jb, m * 7 + jb)
    ref = None
    for pf, rf in itertools.product(FACTS, FACTS):
        P = P0.copy(order="F")
        piv, info = o.pdfact(P, nbmin=nbmin, ndiv=ndiv, pfact=pf, rfact=rf)
        assert info == 0
        if ref is None:
            ref = (P, piv)
            continue
        assert np.array_equal(piv, ref[1]), (pf, rf)
        assert np.array_equal(P, ref[0]), (pf, rf, np.abs(P - ref[0]).max())


@pytest.mark.parametrize("pf,rf", list(itertools.product(FACTS, FACTS)))
def test_variant_pivots_and_factors_equal_lapack(pf, rf):
    m, jb = 400, 128
    P0 = _panel(m, jb, 11)
    P = P0.copy(order="F")
    piv, info = o.pdfact(P, nbmin=16, ndiv=2, pfact=pf, rfact=rf)
    lu, lpiv = sl.lu_factor(P0)
    assert info == 0
    assert np.array_equal(piv, lpiv[:jb])
    assert np.max(np.abs(P - lu[:, :jb])) < 1e-12


def test_solve_with_every_variant_is_the_same():
    n, nb = 500, 64
    ref = o.pdgesv(n, nb, nbmin=8)
    for pf, rf in itertools.product(FACTS, FACTS):
        w = o.pdgesv(n, nb, nbmin=8, pfact=pf, rfact=rf)
        assert w["info"] == 0
        assert np.array_equal(w["ipiv"], ref["ipiv"])
        assert np.array_equal(w["x"], ref["x"])
    assert o.residual(n, ref["x"])["resid"] < 16.0


def test_zero_pivot_is_reported_by_every_variant():
    """Column 2 is zero: every update it receives multiplies one of its own
    zero entries, so its pivot is an exact zero (info = 3) whichever variant
    factors it, and the later columns are still factored."""
    P0 = _panel(60, 16, 3)
    P0[:, 2] = 0.0
    for pf, rf in itertools.product(FACTS, FACTS):
        P = P0.copy(order="F")
        _, info = o.pdfact(P, nbmin=4, ndiv=2, pfact=pf, rfact=rf)
        assert info == 3, (pf, rf)


def test_unknown_variant_is_refused():
    with pytest.raises(ValueError):
        o.pdfact(_panel(8, 4, 1), pfact=300)

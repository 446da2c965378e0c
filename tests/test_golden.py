"""The oracle reproduces the committed fixtures (tests/golden/hpl_golden.npz,
written by tests/golden/make_golden.py with LAPACK cross-checks) bit for
bit: solves, generator spot values of the full-size matrices, and the
leading pivots of the N=193280 matrix."""
import os

import numpy as np
import pytest

from oracle import oracle as o

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "hpl_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


@pytest.mark.parametrize("n,nb", [(64, 16), (257, 32), (1000, 64),
                                  (4096, 256)])
def test_oracle_solves_match_fixture(gold, n, nb):
    r = o.pdgesv(n, nb)
    assert np.array_equal(r["ipiv"], gold["ipiv_%d_%d" % (n, nb)])
    assert np.array_equal(r["x"], gold["x_%d_%d" % (n, nb)])
    assert o.residual(n, r["x"])["resid"] == float(gold["resid_%d_%d" % (n, nb)])


@pytest.mark.parametrize("n", [193280, 131072, 229376])
def test_generator_spot_values(gold, n):
    ij = gold["spot_ij_%d" % n]
    vals = gold["spot_val_%d" % n]
    for (i, j), v in zip(ij, vals):
        assert o.matgen_tile(n, int(i), int(j), 1, 1)[0, 0] == v
    # every value is a generator value: in [-0.5, 0.5]
    assert np.all(np.abs(vals) <= 0.5)


@pytest.mark.parametrize("key,n,k", [
    ("prefix_ipiv_193280_64", 193280, 64),
    ("prefix_ipiv_193280_4096", 193280, 4096),
    ("prefix_ipiv_131072_64", 131072, 64),
    ("prefix_ipiv_131072_4096", 131072, 4096),
    ("prefix_ipiv_229376_64_shift57344", 229376, 64),
    ("prefix_ipiv_229376_4096_shift57344", 229376, 4096)])
def test_prefix_pivots_fixture_is_a_permutation_step(gold, key, n, k):
    p = gold[key]
    assert p.shape == (k,)
    assert np.all(p >= np.arange(k)) and np.all(p < n)
    # the short prefix is the start of the long one
    short = gold[key.replace("_%d" % k, "_64", 1)] if k != 64 else p
    assert np.array_equal(p[:64], short)
    if "shift" in key:  # column diagonal dominance: no interchange
        assert np.array_equal(p, np.arange(k))

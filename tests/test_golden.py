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


def test_prefix_pivots_fixture_is_a_permutation_step(gold):
    p = gold["prefix_ipiv_193280_64"]
    assert p.shape == (64,)
    assert np.all(p >= np.arange(64)) and np.all(p < 193280)

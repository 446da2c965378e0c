"""The HIP solve against the committed fixtures (tests/golden/hpl_golden.npz):
pivots bit-exact, x within the conditioning of each matrix, residual < 16."""
import os

import numpy as np
import pytest

import rochpl_amd as rp

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "hpl_golden.npz")


@pytest.mark.parametrize("n,nb", [(64, 16), (257, 32), (1000, 64),
                                  (4096, 256)])
def test_pdgesv_matches_fixture(ctx, n, nb):
    print("[gpu-case] golden pdgesv n=%d nb=%d" % (n, nb), flush=True)
    gold = np.load(GOLD, allow_pickle=False)
    mat = ctx.pmat_alloc(n, nb)
    try:
        ctx.pdrandmat(n, n + 1, nb, mat.A, mat.ld)
        assert ctx.pdgesv(rp.palg_default(), mat) == 0
        ipiv = ctx.download_vec(mat.ipiv, n, np.int32)
        x = ctx.download_vec(mat.X, n)
        chk = ctx.pdtest_check(mat)
    finally:
        ctx.pmat_free(mat)
    assert np.array_equal(ipiv, gold["ipiv_%d_%d" % (n, nb)])
    xg = gold["x_%d_%d" % (n, nb)]
    # same pivots, different summation order: x differs by ~cond * eps
    assert np.linalg.norm(x - xg) / np.linalg.norm(xg) < 1e-8
    assert chk["resid"] < 16.0

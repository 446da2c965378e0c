"""Writes tests/golden/hpl_golden.npz: fixed vectors for the parity tests.

The reference (rocHPL) holds no golden vectors and cannot run here, so the
fixtures come from the oracle (oracle/hpl_oracle.c), each cross-checked by
this script against an independent computation before it is written:
  * solves (N, NB) of the seeded HPL matrix: pivots and x from the oracle's
    P=Q=1 HPL_pdgesv restatement, required to equal LAPACK dgetrf's pivots
    and to match scipy's lu_solve x to 1e-12; the scaled residual;
  * generator spot values: elements (i, j) of the N=193280 (FP64 solve),
    N=131072 and N=229376 (mixed forms) matrices at their corners and
    random positions, each also computed by exact integer arithmetic here
    (LCG jump by repeated squaring in Python integers);
  * the first 64 pivots of the N=193280 matrix, from LAPACK dgetrf on the
    N x 64 column prefix (the first K pivots of LU(A) are those of
    LU(A[:, :K])), and likewise of the N=131072 matrix and of the N=229376
    matrix with N/4 added to its diagonal (the mixed-precision forms); the
    same for the first 4096 pivots of all three (K = 4096).

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.linalg as sl
import scipy.linalg.lapack as la

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(
    __file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as o  # noqa: E402

MULT = 6364136223846793005
IADD = 1
MASK = (1 << 64) - 1


def lcg_state_exact(seed, k):
    """State k of the sequence: s_0 = MULT*seed + IADD, s_{k+1} = MULT*s_k +
    IADD (mod 2^64), by binary powering of the affine map."""
    a, c = 1, 0            # accumulated map x -> a x + c
    pa, pc = MULT, IADD    # map for 2^bit steps
    kk = k
    while kk:
        if kk & 1:
            a, c = (pa * a) & MASK, (pa * c + pc) & MASK
        pa, pc = (pa * pa) & MASK, (pa * pc + pc) & MASK
        kk >>= 1
    s0 = (MULT * seed + IADD) & MASK
    return (a * s0 + c) & MASK


def main():
    out = {}
    rng = np.random.default_rng(2024)
    # ---- solves
    for (n, nb) in [(64, 16), (257, 32), (1000, 64), (4096, 256)]:
        r = o.pdgesv(n, nb)
        A = o.matgen(n, n)
        b = o.matgen_tile(n, 0, n, n, 1)[:, 0]
        lu, piv = sl.lu_factor(A)
        assert np.array_equal(r["ipiv"], piv), (n, nb)
        xl = sl.lu_solve((lu, piv), b)
        # two backward-stable solves differ by about cond(A) * eps
        kappa = np.linalg.cond(A, 1)
        err = np.linalg.norm(r["x"] - xl) / np.linalg.norm(xl)
        print("n=%d nb=%d: cond_1 %.3g, |x - x_lapack|/|x| %.3g" % (
            n, nb, kappa, err))
        assert err < 10 * kappa * np.finfo(float).eps
        res = o.residual(n, r["x"])
        assert res["resid"] < 16
        out["ipiv_%d_%d" % (n, nb)] = r["ipiv"].astype(np.int32)
        out["x_%d_%d" % (n, nb)] = r["x"]
        out["resid_%d_%d" % (n, nb)] = np.float64(res["resid"])
    # ---- generator spot values of the large matrices
    for n in (193280, 131072, 229376):
        ij = [(0, 0), (n - 1, 0), (0, n - 1), (n - 1, n - 1), (0, n),
              (n - 1, n)]
        ij += [(int(i), int(j)) for i, j in
               zip(rng.integers(0, n, 12), rng.integers(0, n + 1, 12))]
        vals = []
        for (i, j) in ij:
            v = o.matgen_tile(n, i, j, 1, 1)[0, 0]
            s = lcg_state_exact(o.HPL_ISEED, i + j * n)
            assert o.state(o.HPL_ISEED, i + j * n) == s
            vals.append(v)
        out["spot_ij_%d" % n] = np.array(ij, dtype=np.int64)
        out["spot_val_%d" % n] = np.array(vals)
    # ---- leading pivots of the config-2 matrix (LAPACK on the prefix)
    n, k = 193280, 64
    pref = np.asfortranarray(o.matgen_tile(n, 0, 0, n, k))
    _, lpiv, info = la.dgetrf(pref, overwrite_a=1)
    assert info == 0
    out["prefix_ipiv_193280_64"] = lpiv[:k].astype(np.int32)
    # ... and of the mixed-precision forms' matrices (the FP8 form's with
    # N/4 added to the diagonal)
    for n, shift in ((131072, 0.0), (229376, 229376 / 4.0)):
        pref = np.asfortranarray(o.matgen_tile(n, 0, 0, n, k))
        for j in range(k):
            pref[j, j] += shift
        _, lpiv, info = la.dgetrf(pref, overwrite_a=1)
        assert info == 0
        key = "prefix_ipiv_%d_64" % n + ("_shift%d" % shift if shift else "")
        out[key] = lpiv[:k].astype(np.int32)
        if shift:  # column diagonal dominance keeps every pivot on the diagonal
            assert np.array_equal(out[key], np.arange(k))
    # ... and the first 4096 pivots of each (the longer check the bench runs)
    kl = 4096
    for n, shift in ((193280, 0.0), (131072, 0.0), (229376, 229376 / 4.0)):
        pref = np.asfortranarray(o.matgen_tile(n, 0, 0, n, kl))
        for j in range(kl):
            pref[j, j] += shift
        _, lpiv, info = la.dgetrf(pref, overwrite_a=1)
        del pref
        assert info == 0
        key = "prefix_ipiv_%d_%d" % (n, kl) + ("_shift%d" % shift if shift else "")
        out[key] = lpiv[:kl].astype(np.int32)
        short = "prefix_ipiv_%d_64" % n + ("_shift%d" % shift if shift else "")
        assert np.array_equal(out[key][:64], out[short]), key
        print("%s: %d pivots" % (key, kl), flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "hpl_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, sorted(out))


if __name__ == "__main__":
    main()

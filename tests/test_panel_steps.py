"""The step list the persistent panel kernel walks (hplx_panel_steps, recorded
on the host from the library's own recursion; no GPU needed).

Checked two ways: structural properties every correct list has (the leaves
cover the panel's columns once, in order; every operand lies inside the
panel for any panel height), and by running the steps on the CPU with a
small numpy interpreter (a leaf is textbook unblocked LU with partial
pivoting on its columns, rows swapped across the panel width; a solve and a
product are what their names say) and comparing the result with the
oracle's restatement of HPL_pdrpan*T / HPL_pdpan*T for the same variant:
equal pivots, factors to rounding."""
import numpy as np
import pytest

import rochpl_amd as rp
from oracle import oracle as o

_FACT = {"L": rp.HPLX_LEFT_LOOKING, "C": rp.HPLX_CROUT,
         "R": rp.HPLX_RIGHT_LOOKING}


def _palg(nbmin, ndiv, pf="R", rf="R"):
    a = rp.palg_default()
    a.nbmin, a.nbdiv, a.pfact, a.rfact = nbmin, ndiv, _FACT[pf], _FACT[rf]
    return a


def _run_steps(steps, P):
    """Apply the recorded steps to the m x jb panel P (in place); returns the
    panel-relative pivot rows."""
    m, jb = P.shape
    piv = np.full(jb, -1, dtype=np.int64)
    for kind, m_fix, m_sub, n, k, ar, ac, br, bc, cr, cc in steps:
        if kind == rp.PANEL_LEAF:
            for j in range(ac, ac + n):
                p = j + int(np.argmax(np.abs(P[j:, j])))
                piv[j] = p
                if p != j:
                    P[[j, p], :] = P[[p, j], :]
                if P[j, j] != 0.0:
                    P[j + 1:, j] /= P[j, j]
                P[j + 1:, j + 1:ac + n] -= np.outer(P[j + 1:, j],
                                                    P[j, j + 1:ac + n])
        elif kind == rp.PANEL_TRSM:
            L = np.tril(P[ar:ar + k, ac:ac + k], -1) + np.eye(k)
            P[br:br + k, bc:bc + n] = np.linalg.solve(L, P[br:br + k, bc:bc + n])
        else:
            rows = m_fix if m_fix >= 0 else m - m_sub
            P[cr:cr + rows, cc:cc + n] -= (P[ar:ar + rows, ac:ac + k]
                                           @ P[br:br + k, bc:bc + n])
    return piv


@pytest.mark.parametrize("pf", ["L", "C", "R"])
@pytest.mark.parametrize("rf", ["L", "C", "R"])
@pytest.mark.parametrize("jb,nbmin,ndiv", [(96, 16, 2), (100, 16, 3),
                                           (128, 48, 2), (64, 32, 2)])
def test_panel_steps_reproduce_the_oracle(pf, rf, jb, nbmin, ndiv):
    m = 3 * jb + 17
    steps = rp.panel_steps(_palg(nbmin, ndiv, pf, rf), jb)
    P0 = np.asfortranarray(o.matgen_tile(m, 0, 0, m, jb, seed=jb + nbmin))
    P = P0.copy()
    piv = _run_steps(steps, P)
    W = P0.copy(order="F")
    wpiv, winfo = o.pdfact(W, nbmin=nbmin, ndiv=ndiv, pfact=_FACT[pf],
                           rfact=_FACT[rf])
    assert winfo == 0
    assert np.array_equal(piv, wpiv)
    assert np.max(np.abs(P - W)) < 1e-10


@pytest.mark.parametrize("jb", [1, 3, 41, 96, 256, 896, 1000, 1024])
@pytest.mark.parametrize("nbmin,ndiv", [(32, 2), (16, 3), (64, 2), (8, 4)])
@pytest.mark.parametrize("pf,rf", [("R", "R"), ("L", "C"), ("C", "L")])
def test_panel_steps_cover_the_panel(jb, nbmin, ndiv, pf, rf):
    steps = rp.panel_steps(_palg(nbmin, ndiv, pf, rf), jb)
    leaves = steps[steps[:, 0] == rp.PANEL_LEAF]
    # the leaves factor every column once, left to right, <= 32 at a time
    cols = np.concatenate([np.arange(s[6], s[6] + s[3]) for s in leaves])
    assert np.array_equal(cols, np.arange(jb))
    assert leaves[:, 3].max() <= min(32, max(nbmin, 1)) or nbmin > 32
    assert leaves[:, 3].max() <= 32
    # every operand inside the panel whatever its height m >= jb
    for kind, m_fix, m_sub, n, k, ar, ac, br, bc, cr, cc in steps:
        if kind == rp.PANEL_LEAF:
            continue
        assert 0 < n and 0 < k
        assert 0 <= ac and ac + k <= jb and 0 <= bc and bc + n <= jb
        if kind == rp.PANEL_TRSM:
            assert ar == ac and br + k <= jb and ar <= br
            continue
        assert cc == bc and m_fix <= jb
        if m_fix < 0:
            # rows from ar, cr to the panel's last row
            assert m_sub == ar == cr and 0 < m_sub <= jb
        assert br + k <= jb and ar >= 0 and cr >= 0


def test_panel_steps_counts_at_nb_896():
    """NB = 896, NBMIN = 32, NDIV = 2, right-looking: 28 leaves, and one
    solve and one product after each of the 27 leaves but the last."""
    steps = rp.panel_steps(_palg(32, 2), 896)
    kinds = np.bincount(steps[:, 0], minlength=3)
    assert list(kinds) == [28, 27, 27]
    assert steps[-1, 0] == rp.PANEL_LEAF


def test_panel_steps_rejects_bad_arguments():
    with pytest.raises(rp.HplxError):
        rp.panel_steps(_palg(32, 2), 0)
    with pytest.raises(rp.HplxError):
        rp.panel_steps(_palg(32, 2), 1025)
    a = _palg(32, 2)
    a.pfact = 7
    with pytest.raises(rp.HplxError):
        rp.panel_steps(a, 64)

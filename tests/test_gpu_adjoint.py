"""Differentiable multi-RHS solve X = A^{-1} B through the C ABI:
hplx_pdgetrf / hplx_pdgetrs / hplx_pdgesv_adjoint.

The reference has no adjoint; its parity anchors are the textbook identities
  grad_B = A^{-T} G,  grad_A = -grad_B X^T
checked against (1) LAPACK (scipy) FP64 solves on the CPU and (2) central
finite differences of loss = sum(G * X), h = 1e-6, on 32 random entries
(the config: N = 16384, A = U(-0.5,0.5) + 2 sqrt(N) I from the HPL generator,
B and G standard normal from numpy's PCG64 with the seeds below)."""
import os
import subprocess
import sys

import numpy as np
import pytest
import scipy.linalg as sl

import rochpl_amd as rp
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-case] " + msg, flush=True)


class Solver:
    """Factors of A held on the device by the library."""

    def __init__(self, ctx, n, nb, shift, seed=rp.HPL_ISEED):
        self.ctx, self.n, self.nb, self.shift, self.seed = ctx, n, nb, shift, seed
        self.mat = ctx.pmat_alloc(n, nb)

    def factor(self, perturb=None):
        c, m = self.ctx, self.mat
        c.pdrandmat(self.n, self.n, self.nb, m.A, m.ld, self.seed, self.shift)
        if perturb is not None:
            (i, j, h) = perturb
            ptr = m.A + 8 * (i + j * m.ld)
            cur = c.download_vec(ptr, 1)
            cur[0] += h
            rp.lib().hplx_memcpy_h2d(c.h, rp.ctypes.c_void_p(ptr),
                                     cur.ctypes.data_as(rp.ctypes.c_void_p), 8)
        info = c.pdgetrf(rp.palg_default(), m)
        assert info == 0

    def solve(self, Bhost, trans=0):
        n, k = Bhost.shape
        d = self.ctx.alloc(8 * n * k)
        d.upload(np.asfortranarray(Bhost))
        self.ctx.pdgetrs(self.mat, trans, k, d.ptr, n)
        out = d.download((n, k), np.float64)
        d.free()
        return out

    def free(self):
        self.ctx.pmat_free(self.mat)


# (2100, 1024, 513): the transposed solve's scratch (2^20 doubles) holds
# one 1024 x 513 partial product only, so the L21^T z steps of the two
# 1024-wide panels run unsplit and the 52-wide last panel's split over K;
# the other cases split over K throughout.  (1300, 256,
# 40): 40 right-hand sides (a partial 64-wide tile) and a 20-wide last panel
# (K = 20 in its products).
@pytest.mark.parametrize("n,nb,shift,nrhs", [(1000, 64, 0.0, 5),
                                             (1500, 256, 2 * 1500 ** 0.5, 7),
                                             (777, 96, 0.0, 1),
                                             (2100, 1024, 0.0, 513),
                                             (1300, 256, 0.0, 40)])
def test_getrs_matches_lapack(ctx, n, nb, shift, nrhs):
    say("getrs n=%d nb=%d" % (n, nb))
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, nrhs))
    s = Solver(ctx, n, nb, shift)
    try:
        s.factor()
        ipiv = ctx.download_vec(s.mat.ipiv, n, np.int32)
        X = s.solve(B, 0)
        Z = s.solve(B, 1)
    finally:
        s.free()
    A = o.matgen(n, n, diag_shift=shift)
    lu, piv = sl.lu_factor(A)
    assert np.array_equal(ipiv, piv)
    Xr = sl.lu_solve((lu, piv), B)
    Zr = sl.lu_solve((lu, piv), B, trans=1)
    cond = np.linalg.cond(A, 1)
    tol = 1e-15 * cond * 10 * n ** 0.5
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) < tol
    assert np.linalg.norm(Z - Zr) / np.linalg.norm(Zr) < tol
    # and they really solve the systems
    assert np.linalg.norm(A @ X - B) / np.linalg.norm(B) < tol
    assert np.linalg.norm(A.T @ Z - B) / np.linalg.norm(B) < tol


def test_adjoint_config5_vs_cpu_and_finite_differences(ctx):
    n, nb, nrhs = 16384, 512, 64
    shift = 2.0 * n ** 0.5
    say("adjoint N=%d NB=%d nrhs=%d" % (n, nb, nrhs))
    rng = np.random.default_rng(20240501)
    B = rng.standard_normal((n, nrhs))
    G = rng.standard_normal((n, nrhs))
    s = Solver(ctx, n, nb, shift)
    try:
        s.factor()
        X = s.solve(B)
        dX = ctx.alloc(8 * n * nrhs)
        dX.upload(np.asfortranarray(X))
        dG = ctx.alloc(8 * n * nrhs)
        dG.upload(np.asfortranarray(G))
        dgB = ctx.alloc(8 * n * nrhs)
        dgA = ctx.alloc(8 * n * n)
        ctx.pdgesv_adjoint(s.mat, nrhs, dX.ptr, n, dG.ptr, n, dgB.ptr, n,
                           dgA.ptr, n)
        gB = dgB.download((n, nrhs), np.float64)
        rows = rng.choice(n, 64, replace=False)
        gA = dgA.download((n, n), np.float64)
        for d in (dX, dG, dgB, dgA):
            d.free()

        # ---- CPU FP64 reference gradient (LAPACK)
        A = o.matgen(n, n, diag_shift=shift)
        lu = sl.lu_factor(A)
        Xr = sl.lu_solve(lu, B)
        gBr = sl.lu_solve(lu, G, trans=1)
        assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-9
        assert np.linalg.norm(gB - gBr) / np.linalg.norm(gBr) <= 1e-9
        gAr_rows = -gBr[rows] @ Xr.T
        assert np.linalg.norm(gA[rows] - gAr_rows) / \
            np.linalg.norm(gAr_rows) <= 1e-9
        # grad_A is exactly the rank-nrhs product of what the library returned
        assert np.linalg.norm(gA[rows] + gB[rows] @ X.T) <= \
            1e-12 * np.linalg.norm(gA[rows])
        del lu

        # ---- central finite differences of loss = sum(G * X), h = 1e-6.
        # L(A + dA) - L(A) = sum(G * delta) with delta = X(A + dA) - X, the
        # solution of (A + dA) delta = -dA X: delta is O(h) and is computed
        # with the perturbed factors directly, instead of as the difference
        # of two O(1) solutions (which would cancel all but ~7 digits).
        h = 1e-6
        fd, gsel, what = [], [], []
        for e in range(32):
            i, j = int(rng.integers(n)), int(rng.integers(n))
            a = float(A[i, j])
            hp = (a + h) - a          # the perturbations as stored, exactly
            hm = a - (a - h)
            R = np.zeros((n, nrhs))
            s.factor((i, j, +h))
            R[i, :] = -hp * X[j, :]
            dp = s.solve(R)
            s.factor((i, j, -h))
            R[i, :] = hm * X[j, :]
            dm = s.solve(R)
            fd.append(float(np.sum(G * (dp - dm))) / (hp + hm))
            gsel.append(gA[i, j])
            what.append("A[%d,%d]" % (i, j))
        # B enters linearly: delta = A^{-1} (h e_i e_k^T), one column per
        # entry, all 32 in one solve with the unperturbed factors
        s.factor()
        ik = [(int(rng.integers(n)), int(rng.integers(nrhs)))
              for _ in range(32)]
        R = np.zeros((n, 32))
        for e, (i, k) in enumerate(ik):
            R[i, e] = h
        D = s.solve(R)
        for e, (i, k) in enumerate(ik):
            fd.append(float(np.dot(G[:, k], D[:, e])) / h)
            gsel.append(gB[i, k])
            what.append("B[%d,%d]" % (i, k))
    finally:
        s.free()
    fd, gsel = np.array(fd), np.array(gsel)
    rel = np.abs(fd - gsel) / np.abs(gsel)
    worst = int(np.argmax(rel))
    print("[adjoint] central FD (h=1e-6) vs the adjoint: per-entry rel err "
          "max %.3e (%s) median %.3e over %d entries" %
          (rel.max(), what[worst], np.median(rel), len(rel)), flush=True)
    assert np.all(rel <= 1e-9), [(what[e], rel[e]) for e in range(len(rel))
                                 if rel[e] > 1e-9]


def test_autograd_function_wrapper():
    """The test-only torch.autograd.Function (tests/autograd_case.py) against
    torch.linalg.solve's FP64 gradient.  It runs in a child process that
    imports torch before loading libhplx.so, so that both use torch's HIP
    runtime (this process has already loaded the system one)."""
    pytest.importorskip("torch")
    say("autograd wrapper n=2048 (child process)")
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, "-u",
                        os.path.join(here, "autograd_case.py")],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0
    assert "AUTOGRAD-OK" in r.stdout

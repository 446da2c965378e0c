"""Differentiable multi-RHS solve X = A^{-1} B through the C ABI:
hplx_pdgetrf / hplx_pdgetrs / hplx_pdgesv_adjoint.

The reference has no adjoint; its parity anchors are the textbook identities
  grad_B = A^{-T} G,  grad_A = -grad_B X^T
checked against (1) LAPACK (scipy) FP64 solves on the CPU and (2) central
finite differences of loss = sum(G * X), h = 1e-6, on 32 random entries
(the config: N = 16384, A = U(-0.5,0.5) + 2 sqrt(N) I from the HPL generator,
B and G standard normal from numpy's PCG64 with the seeds below)."""
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


@pytest.mark.parametrize("n,nb,shift,nrhs", [(1000, 64, 0.0, 5),
                                             (1500, 256, 2 * 1500 ** 0.5, 7),
                                             (777, 96, 0.0, 1)])
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
        del A, lu

        # ---- central finite differences of loss = sum(G * X), h = 1e-6
        h = 1e-6

        def loss(Xv):
            return float(np.sum(G * Xv))

        fd_rel = []
        for e in range(32):
            i, j = int(rng.integers(n)), int(rng.integers(n))
            s.factor((i, j, +h))
            lp = loss(s.solve(B))
            s.factor((i, j, -h))
            lm = loss(s.solve(B))
            fd = (lp - lm) / (2 * h)
            fd_rel.append(abs(fd - gA[i, j]) / abs(gA[i, j]))
        s.factor()
        for e in range(32):
            i, k = int(rng.integers(n)), int(rng.integers(nrhs))
            Bp = B.copy()
            Bp[i, k] += h
            lp = loss(s.solve(Bp))
            Bp[i, k] -= 2 * h
            lm = loss(s.solve(Bp))
            fd = (lp - lm) / (2 * h)
            fd_rel.append(abs(fd - gB[i, k]) / abs(gB[i, k]))
    finally:
        s.free()
    fd_rel = np.array(fd_rel)
    print("[adjoint] FD relative error: max %.3e median %.3e (A entries max "
          "%.3e, B entries max %.3e)" % (fd_rel.max(), np.median(fd_rel),
                                         fd_rel[:32].max(), fd_rel[32:].max()),
          flush=True)
    assert np.median(fd_rel) <= 1e-7
    assert fd_rel.max() <= 1e-5


def test_autograd_function_wrapper(ctx):
    """A test-only torch.autograd.Function around the C ABI: forward factors
    and solves, backward calls hplx_pdgesv_adjoint; compared with torch's own
    FP64 linear solve gradient."""
    torch = pytest.importorskip("torch")
    n, nb, nrhs = 2048, 256, 16
    say("autograd wrapper n=%d" % n)
    dev = torch.device("cuda:0")

    class LUSolve(torch.autograd.Function):
        @staticmethod
        def forward(fctx, A, B):
            mat = ctx.pmat_alloc(n, nb)
            Ah = A.detach().double().cpu().numpy()
            buf = np.zeros((mat.ld, n + 1), order="F")
            buf[:n, :n] = Ah
            rp.lib().hplx_memcpy_h2d(ctx.h, rp.ctypes.c_void_p(mat.A),
                                     buf.ctypes.data_as(rp.ctypes.c_void_p),
                                     buf.nbytes)
            assert ctx.pdgetrf(rp.palg_default(), mat) == 0
            Xd = B.detach().double().t().contiguous().t().clone()  # column-major
            torch.cuda.synchronize()
            ctx.pdgetrs(mat, 0, nrhs, Xd.data_ptr(), n)
            fctx.mat = mat
            fctx.save_for_backward(Xd)
            return Xd.clone()

        @staticmethod
        def backward(fctx, Gout):
            (Xd,) = fctx.saved_tensors
            Gc = Gout.double().t().contiguous().t()
            gB = torch.empty_like(Gc).t().contiguous().t()
            gA = torch.empty((n, n), dtype=torch.float64,
                             device=dev).t().contiguous().t()
            torch.cuda.synchronize()
            ctx.pdgesv_adjoint(fctx.mat, nrhs, Xd.data_ptr(), n, Gc.data_ptr(),
                               n, gB.data_ptr(), n, gA.data_ptr(), n)
            ctx.pmat_free(fctx.mat)
            return gA, gB

    An = o.matgen(n, n, diag_shift=2 * n ** 0.5)
    A1 = torch.tensor(An, device=dev, requires_grad=True)
    g = torch.Generator(device="cpu").manual_seed(7)
    Bt = torch.randn(n, nrhs, generator=g, dtype=torch.float64).to(dev)
    Gt = torch.randn(n, nrhs, generator=g, dtype=torch.float64).to(dev)
    B1 = Bt.clone().requires_grad_(True)
    X1 = LUSolve.apply(A1, B1)
    (X1 * Gt).sum().backward()
    A2 = torch.tensor(An, device=dev, requires_grad=True)
    B2 = Bt.clone().requires_grad_(True)
    X2 = torch.linalg.solve(A2, B2)
    (X2 * Gt).sum().backward()
    torch.cuda.synchronize()

    def rel(a, b):
        return (torch.linalg.norm(a - b) / torch.linalg.norm(b)).item()

    assert rel(X1, X2) < 1e-12
    assert rel(B1.grad, B2.grad) < 1e-12
    assert rel(A1.grad, A2.grad) < 1e-12

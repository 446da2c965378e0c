"""Child process of tests/test_gpu_adjoint.py::test_autograd_function_wrapper.

A test-only torch.autograd.Function around the C ABI: forward factors A
(hplx_pdgetrf) and solves (hplx_pdgetrs), backward calls hplx_pdgesv_adjoint;
the result is compared with torch's own FP64 linear-solve gradient.

torch ships its own HIP runtime (torch/lib/libamdhip64.so).  torch is imported
BEFORE libhplx.so is loaded so that libhplx binds to that same runtime (both
carry the soname libamdhip64.so.7): torch's streams and device pointers then
belong to the runtime the library launches on.  The parent test process has
already loaded the system runtime, hence the separate process.

All torch work and all library work go to one torch stream (set on the
library with hplx_set_stream), so they are ordered without extra syncs.
Prints the three relative errors, then "AUTOGRAD-OK" and exits 0 when all are
below 1e-12; exits non-zero otherwise."""
import os
import sys

import torch  # must precede the first load of libhplx.so

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import numpy as np  # noqa: E402

import rochpl_amd as rp  # noqa: E402
from oracle import oracle as o  # noqa: E402


def main(n=2048, nb=256, nrhs=16):
    dev = torch.device("cuda:0")
    torch.zeros(1, device=dev)  # initialise torch's device context first
    ctx = rp.Context(0)
    stream = torch.cuda.Stream()  # default priority
    ctx.set_stream(stream.cuda_stream)  # library work queues behind torch's
    with torch.cuda.stream(stream):
        r = run(ctx, dev, n, nb, nrhs)
    torch.cuda.synchronize()
    ctx.close()
    print("AUTOGRAD x=%.3e gB=%.3e gA=%.3e" % r, flush=True)
    if max(r) >= 1e-12:
        sys.exit(1)
    print("AUTOGRAD-OK", flush=True)


def run(ctx, dev, n, nb, nrhs):
    class LUSolve(torch.autograd.Function):
        @staticmethod
        def forward(fctx, A, B):
            mat = ctx.pmat_alloc(n, nb)
            buf = np.zeros((mat.ld, n + 1), order="F")
            buf[:n, :n] = A.detach().double().cpu().numpy()
            rp.lib().hplx_memcpy_h2d(ctx.h, rp.ctypes.c_void_p(mat.A),
                                     buf.ctypes.data_as(rp.ctypes.c_void_p),
                                     buf.nbytes)
            assert ctx.pdgetrf(rp.palg_default(), mat) == 0
            Xd = B.detach().double().t().contiguous().t().clone()  # column-major
            ctx.pdgetrs(mat, 0, B.shape[1], Xd.data_ptr(), n)
            fctx.mat = mat
            fctx.save_for_backward(Xd)
            return Xd.clone()

        @staticmethod
        def backward(fctx, Gout):
            (Xd,) = fctx.saved_tensors
            k = Xd.shape[1]
            Gc = Gout.double().t().contiguous().t()
            gB = torch.empty_like(Gc).t().contiguous().t()
            gA = torch.empty((n, n), dtype=torch.float64,
                             device=dev).t().contiguous().t()
            ctx.pdgesv_adjoint(fctx.mat, k, Xd.data_ptr(), n, Gc.data_ptr(),
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

    return (rel(X1, X2), rel(B1.grad, B2.grad), rel(A1.grad, A2.grad))


if __name__ == "__main__":
    main()

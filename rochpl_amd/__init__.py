"""rochpl_amd -- MI355X-native drop-in for rocHPL's P=Q=1 HPL_pdgesv path.

The product is the C-ABI library ``rochpl_amd/libhplx.so`` (hand-written HIP
for gfx950, declared in ``include/hplx.h``).  This module is a thin ctypes
mirror of that ABI with the reference's names (HPL_pdgesv -> pdgesv, ...), so
Python callers and tests read like the reference's own driver
(src/HPL_pdtest.cpp).  There is no CPU fallback: every call goes to the HIP
library and fails loudly if it is missing.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhplx.so")
CSRC = os.path.join(_HERE, "csrc")
HPL_ISEED = 100  # include/hpl_ptest.hpp:57

HPLX_SUCCESS = 0
HPLX_ERR_INVALID = -1
HPLX_ERR_UNSUPPORTED = -2
HPLX_ERR_HIP = -3
HPLX_ERR_ALLOC = -4

HPLX_LEFT_LOOKING, HPLX_CROUT, HPLX_RIGHT_LOOKING = 301, 302, 303
HPLX_1RING, HPLX_1RING_M, HPLX_2RING, HPLX_2RING_M, HPLX_BLONG, \
    HPLX_BLONG_M = 401, 402, 403, 404, 405, 406
HPLX_SWAP00, HPLX_SWAP01, HPLX_SW_MIX, HPLX_NO_SWP = 451, 452, 453, 499
# mixed precision (no counterpart in the reference)
HPLX_STORE_FP32, HPLX_STORE_FP16 = 32, 16
HPLX_UPDATE_FP32, HPLX_UPDATE_BF16, HPLX_UPDATE_FP8 = 0, 1, 2
HPLX_OPT_LOOKAHEAD = 1
HPLX_OPT_PANEL_KERNEL = 2

# Every symbol include/hplx.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "hplx_palg_default", "hplx_palg_check", "hplx_ctx_create", "hplx_ctx_destroy",
    "hplx_last_error", "hplx_set_stream", "hplx_synchronize", "hplx_num_cus",
    "hplx_set_option", "hplx_get_option", "hplx_panel_kernel_runs",
    "hplx_panel_steps",
    "hplx_malloc", "hplx_free", "hplx_memcpy_h2d", "hplx_memcpy_d2h",
    "hplx_pmat_alloc", "hplx_pmat_free", "hplx_pdrandmat", "hplx_pdgesv",
    "hplx_pdtest_check", "hplx_pdfact", "hplx_pdfact_palg", "hplx_swap_timing",
    "hplx_panel_timing",
    "hplx_pipid", "hplx_dlaswp00N",
    "hplx_pdupdate", "hplx_dgemm", "hplx_pdtrsv", "hplx_update_stats",
    "hplx_update_timing",
    "hplx_pdgetrf", "hplx_pdgetrs", "hplx_pdgetrs_noalloc", "hplx_pdgesv_adjoint",
    "hplx_mxp_opts_default", "hplx_mxp_alloc", "hplx_mxp_free",
    "hplx_mxp_matgen", "hplx_mxp_pdgetrf", "hplx_mxp_pdgetrs",
    "hplx_mxp_pdgesv", "hplx_mxp_residual", "hplx_mxp_update", "hplx_spdfact",
]


class HplxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("hplx error %d: %s" % (code, msg))
        self.code = code


class PAlg(ctypes.Structure):
    """hplx_palg, the mirror of HPL_T_palg (include/hpl_pgesv.hpp:63-81)."""
    _fields_ = [("btopo", ctypes.c_int), ("depth", ctypes.c_int),
                ("nbdiv", ctypes.c_int), ("nbmin", ctypes.c_int),
                ("pfact", ctypes.c_int), ("rfact", ctypes.c_int),
                ("fswap", ctypes.c_int), ("fsthr", ctypes.c_int),
                ("equil", ctypes.c_int), ("L1notran", ctypes.c_int),
                ("Unotran", ctypes.c_int), ("align", ctypes.c_int),
                ("frac", ctypes.c_double), ("its", ctypes.c_int)]


class PMat(ctypes.Structure):
    """hplx_pmat, the mirror of HPL_T_pmat (include/hpl_pgesv.hpp:83-109)."""
    _fields_ = [("A", ctypes.c_void_p), ("X", ctypes.c_void_p),
                ("ipiv", ctypes.c_void_p), ("n", ctypes.c_int),
                ("nb", ctypes.c_int), ("ld", ctypes.c_int),
                ("mp", ctypes.c_int), ("nq", ctypes.c_int),
                ("info", ctypes.c_int), ("ws", ctypes.c_void_p)]


class MxpOpts(ctypes.Structure):
    """hplx_mxp_opts (mixed-precision options)."""
    _fields_ = [("update", ctypes.c_int), ("terms", ctypes.c_int),
                ("max_its", ctypes.c_int), ("solve_fp64", ctypes.c_int),
                ("threshold", ctypes.c_double)]


class MxpPMat(ctypes.Structure):
    """hplx_mxp_pmat: low-precision factors, FP64 solution, pivots."""
    _fields_ = [("A", ctypes.c_void_p), ("X", ctypes.c_void_p),
                ("ipiv", ctypes.c_void_p), ("n", ctypes.c_int),
                ("nb", ctypes.c_int), ("ld", ctypes.c_int),
                ("storage", ctypes.c_int), ("seed", ctypes.c_int),
                ("info", ctypes.c_int), ("diag_shift", ctypes.c_double),
                ("ws", ctypes.c_void_p)]


class MxpResult(ctypes.Structure):
    """hplx_mxp_result: the refinement record."""
    _fields_ = [("its", ctypes.c_int), ("converged", ctypes.c_int),
                ("resid", ctypes.c_double), ("factor_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("update_ms", ctypes.c_double),
                ("update_flops", ctypes.c_double),
                ("update_launches", ctypes.c_int), ("pad", ctypes.c_int),
                ("history", ctypes.c_double * 65)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_
             if k not in ("history", "pad")}
        d["history"] = list(self.history[:self.its + 1])
        return d


_lib = None


def build(force: bool = False) -> str:
    """Compile libhplx.so with hipcc for gfx950 (rochpl_amd/csrc/Makefile)."""
    args = ["make", "-s", "-j8"]
    if force:
        args.append("-B")
    subprocess.run(args, cwd=CSRC, check=True)
    return LIB_PATH


def lib():
    """Load libhplx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("rochpl_amd: %s is missing -- run "
                           "__graft_entry__.build() (no CPU fallback)"
                           % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    i32, f64, vp, sz = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, \
        ctypes.c_size_t
    P = ctypes.POINTER
    sig = {
        "hplx_palg_default": (None, [P(PAlg)]),
        "hplx_palg_check": (i32, [P(PAlg)]),
        "hplx_ctx_create": (i32, [i32, P(vp)]),
        "hplx_ctx_destroy": (i32, [vp]),
        "hplx_last_error": (ctypes.c_char_p, [vp]),
        "hplx_set_stream": (i32, [vp, vp]),
        "hplx_synchronize": (i32, [vp]),
        "hplx_num_cus": (i32, [vp]),
        "hplx_set_option": (i32, [vp, i32, i32]),
        "hplx_get_option": (i32, [vp, i32, P(i32)]),
        "hplx_panel_kernel_runs": (i32, [vp]),
        "hplx_panel_steps": (i32, [P(PAlg), i32, vp, i32]),
        "hplx_malloc": (i32, [vp, sz, P(vp)]),
        "hplx_free": (i32, [vp, vp]),
        "hplx_memcpy_h2d": (i32, [vp, vp, vp, sz]),
        "hplx_memcpy_d2h": (i32, [vp, vp, vp, sz]),
        "hplx_pmat_alloc": (i32, [vp, i32, i32, P(PMat)]),
        "hplx_pmat_free": (i32, [vp, P(PMat)]),
        "hplx_pdrandmat": (i32, [vp, i32, i32, i32, vp, i32, i32, f64]),
        "hplx_pdgesv": (i32, [vp, P(PAlg), P(PMat)]),
        "hplx_pdtest_check": (i32, [vp, P(PMat), i32, f64, P(f64)]),
        "hplx_pdfact": (i32, [vp, i32, i32, vp, i32, i32, i32, vp, i32]),
        "hplx_pdfact_palg": (i32, [vp, vp, vp, i32, i32, i32, vp, i32]),
        "hplx_swap_timing": (i32, [vp, P(f64), P(i32), P(f64)]),
        "hplx_panel_timing": (i32, [vp, P(f64), P(i32), P(f64)]),
        "hplx_pipid": (i32, [vp, i32, vp, vp]),
        "hplx_dlaswp00N": (i32, [vp, i32, i32, vp, i32, vp]),
        "hplx_pdupdate": (i32, [vp, vp, i32, i32, i32, i32, vp, vp]),
        "hplx_dgemm": (i32, [vp, i32, i32, i32, f64, vp, i32, vp, i32, f64,
                             vp, i32]),
        "hplx_pdtrsv": (i32, [vp, i32, i32, vp, i32, vp]),
        "hplx_update_stats": (i32, [vp, P(f64), P(i32), P(f64)]),
        "hplx_update_timing": (i32, [vp, i32, P(f64), P(i32), P(f64)]),
        "hplx_pdgetrf": (i32, [vp, P(PAlg), P(PMat)]),
        "hplx_pdgetrs": (i32, [vp, P(PMat), i32, i32, vp, i32]),
        "hplx_pdgetrs_noalloc": (i32, [P(PMat), i32, i32, vp, i32, vp]),
        "hplx_pdgesv_adjoint": (i32, [vp, P(PMat), i32, vp, i32, vp, i32, vp,
                                      i32, vp, i32]),
        "hplx_mxp_opts_default": (None, [P(MxpOpts)]),
        "hplx_mxp_alloc": (i32, [vp, i32, i32, i32, P(MxpPMat)]),
        "hplx_mxp_free": (i32, [vp, P(MxpPMat)]),
        "hplx_mxp_matgen": (i32, [vp, P(MxpPMat), i32, f64]),
        "hplx_mxp_pdgetrf": (i32, [vp, P(PAlg), P(MxpOpts), P(MxpPMat)]),
        "hplx_mxp_pdgetrs": (i32, [vp, P(MxpPMat), vp]),
        "hplx_mxp_pdgesv": (i32, [vp, P(PAlg), P(MxpOpts), P(MxpPMat),
                                  P(MxpResult)]),
        "hplx_mxp_residual": (i32, [vp, P(MxpPMat), vp, P(f64)]),
        "hplx_mxp_update": (i32, [vp, i32, i32, i32, i32, i32, i32, vp, i32,
                                  vp, i32, vp, i32]),
        "hplx_spdfact": (i32, [vp, i32, i32, vp, i32, i32, i32, vp, i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def palg_default() -> PAlg:
    p = PAlg()
    lib().hplx_palg_default(ctypes.byref(p))
    return p


PANEL_LEAF, PANEL_TRSM, PANEL_GEMM = 0, 1, 2


def panel_steps(palg: PAlg, jb: int):
    """The steps the persistent panel kernel walks for a panel of width jb
    (hplx_panel_steps): an (n, 11) int array of (kind, m_fix, m_sub, n, k,
    ar, ac, br, bc, cr, cc).  Host only, no device needed."""
    import numpy as np
    n = lib().hplx_panel_steps(ctypes.byref(palg), jb, None, 0)
    if n < 0:
        raise HplxError(n, "hplx_panel_steps: invalid arguments")
    out = np.zeros((max(n, 1), 11), dtype=np.int32)
    got = lib().hplx_panel_steps(ctypes.byref(palg), jb,
                                 out.ctypes.data_as(ctypes.c_void_p), n)
    if got != n:
        raise HplxError(min(got, -1), "hplx_panel_steps: changed size")
    return out[:n]


def mxp_opts_default(update: int | None = None,
                     terms: int | None = None) -> MxpOpts:
    o = MxpOpts()
    lib().hplx_mxp_opts_default(ctypes.byref(o))
    if update is not None:
        o.update = update
    if terms is not None:
        o.terms = terms
    return o


def hpl_ld(n: int) -> int:
    """src/matgen/HPL_pdmatgen.cpp:109-112."""
    return ((max(n, 0) + 95) // 128) * 128 + 32


class DeviceArray:
    """A device allocation made through the library (hplx_malloc)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx._chk(lib().hplx_malloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def offset(self, nbytes: int) -> int:
        return self.ptr + int(nbytes)

    def upload(self, arr: np.ndarray, byte_offset: int = 0) -> None:
        arr = np.asarray(arr)
        if not (arr.flags.c_contiguous or arr.flags.f_contiguous):
            arr = np.ascontiguousarray(arr)
        assert byte_offset + arr.nbytes <= self.nbytes
        self.ctx._chk(lib().hplx_memcpy_h2d(
            self.ctx.h, ctypes.c_void_p(self.ptr + byte_offset),
            arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes))

    def download(self, shape, dtype, order="F", byte_offset: int = 0):
        out = np.empty(shape, dtype=dtype, order=order)
        assert byte_offset + out.nbytes <= self.nbytes
        self.ctx._chk(lib().hplx_memcpy_d2h(
            self.ctx.h, out.ctypes.data_as(ctypes.c_void_p),
            ctypes.c_void_p(self.ptr + byte_offset), out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            if self.ctx.h:
                lib().hplx_free(self.ctx.h, ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        # No HIP call while the interpreter shuts down: the HIP runtime (and
        # a profiler attached to it) may already be finalised by then, and
        # the process is about to release the memory anyway.
        if sys.is_finalizing():
            return
        try:
            self.free()
        except Exception:
            pass


class Context:
    """hplx_ctx: device, stream and the solver's small workspace
    (replaces HPL_InitGPU, src/HPL_InitGPU.cpp:29-130)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().hplx_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise HplxError(rc, "hplx_ctx_create(device=%d) failed" % device)
        self.h = h

    def _chk(self, rc: int) -> int:
        if rc < 0:
            raise HplxError(rc, lib().hplx_last_error(self.h).decode())
        return rc

    def close(self) -> None:
        if self.h:
            lib().hplx_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def num_cus(self) -> int:
        return lib().hplx_num_cus(self.h)

    def synchronize(self) -> None:
        self._chk(lib().hplx_synchronize(self.h))

    def set_stream(self, stream_handle: int | None) -> None:
        self._chk(lib().hplx_set_stream(self.h, stream_handle))

    @property
    def lookahead(self) -> bool:
        """HPLX_OPT_LOOKAHEAD: panel k+1 factored beside update k."""
        v = ctypes.c_int()
        self._chk(lib().hplx_get_option(self.h, HPLX_OPT_LOOKAHEAD,
                                        ctypes.byref(v)))
        return bool(v.value)

    @lookahead.setter
    def lookahead(self, on: bool) -> None:
        self._chk(lib().hplx_set_option(self.h, HPLX_OPT_LOOKAHEAD,
                                        1 if on else 0))

    @property
    def panel_kernel(self) -> int:
        """HPLX_OPT_PANEL_KERNEL: 1 every FP64 panel as one persistent
        kernel, 0 none (launch sequence), 2 panels of >= 65536 rows."""
        v = ctypes.c_int()
        self._chk(lib().hplx_get_option(self.h, HPLX_OPT_PANEL_KERNEL,
                                        ctypes.byref(v)))
        return v.value

    @panel_kernel.setter
    def panel_kernel(self, mode) -> None:
        self._chk(lib().hplx_set_option(self.h, HPLX_OPT_PANEL_KERNEL,
                                        int(mode)))

    def panel_kernel_runs(self) -> int:
        """Panels of the last factorisation that ran as the persistent
        panel kernel (hplx_panel_kernel_runs)."""
        n = lib().hplx_panel_kernel_runs(self.h)
        self._chk(min(n, 0))
        return n

    def alloc(self, nbytes: int) -> DeviceArray:
        return DeviceArray(self, nbytes)

    # ---- HPL_pdmatgen / HPL_pdmatfree -------------------------------------
    def pmat_alloc(self, n: int, nb: int) -> PMat:
        m = PMat()
        self._chk(lib().hplx_pmat_alloc(self.h, n, nb, ctypes.byref(m)))
        return m

    def pmat_free(self, mat: PMat) -> None:
        self._chk(lib().hplx_pmat_free(self.h, ctypes.byref(mat)))

    # ---- HPL_pdrandmat ------------------------------------------------------
    def pdrandmat(self, m: int, n: int, nb: int, A: int, lda: int,
                  seed: int = HPL_ISEED, diag_shift: float = 0.0) -> None:
        self._chk(lib().hplx_pdrandmat(self.h, m, n, nb, ctypes.c_void_p(A),
                                       lda, seed, float(diag_shift)))

    # ---- HPL_pdgesv ---------------------------------------------------------
    def pdgesv(self, palg: PAlg, mat: PMat) -> int:
        """Returns 0 or the 1-based column of the first exact zero pivot."""
        return self._chk(lib().hplx_pdgesv(self.h, ctypes.byref(palg),
                                           ctypes.byref(mat)))

    # ---- factor / solve / adjoint (new entry points) ---------------------------
    def pdgetrf(self, palg: PAlg, mat: PMat) -> int:
        return self._chk(lib().hplx_pdgetrf(self.h, ctypes.byref(palg),
                                            ctypes.byref(mat)))

    def pdgetrs(self, mat: PMat, trans: int, nrhs: int, B: int,
                ldb: int) -> None:
        self._chk(lib().hplx_pdgetrs(self.h, ctypes.byref(mat), trans, nrhs,
                                     ctypes.c_void_p(B), ldb))

    def pdgesv_adjoint(self, mat: PMat, nrhs: int, X: int, ldx: int, G: int,
                       ldg: int, gradB: int, ldgb: int, gradA: int | None,
                       ldga: int) -> None:
        self._chk(lib().hplx_pdgesv_adjoint(
            self.h, ctypes.byref(mat), nrhs, ctypes.c_void_p(X), ldx,
            ctypes.c_void_p(G), ldg, ctypes.c_void_p(gradB), ldgb,
            ctypes.c_void_p(gradA) if gradA else None, ldga))

    # ---- HPL_pdtest verification --------------------------------------------
    def pdtest_check(self, mat: PMat, seed: int = HPL_ISEED,
                     diag_shift: float = 0.0) -> dict:
        out = (ctypes.c_double * 9)()
        self._chk(lib().hplx_pdtest_check(self.h, ctypes.byref(mat), seed,
                                          float(diag_shift), out))
        keys = ["resid", "rnorm", "resid_r0_form", "AnormI", "Anorm1",
                "XnormI", "Xnorm1", "BnormI", "r0"]
        return dict(zip(keys, list(out)))

    # ---- components -----------------------------------------------------------
    def pdfact(self, P: int, lda: int, m: int, jb: int, ipiv: int,
               nbmin: int = 32, nbdiv: int = 2, row_offset: int = 0) -> int:
        return self._chk(lib().hplx_pdfact(self.h, nbmin, nbdiv,
                                           ctypes.c_void_p(P), lda, m, jb,
                                           ctypes.c_void_p(ipiv), row_offset))

    def pdfact_palg(self, palg: "PAlg", P: int, lda: int, m: int, jb: int,
                    ipiv: int, row_offset: int = 0) -> int:
        """HPL_pdfact with palg's pfact/rfact variants, nbmin and nbdiv."""
        return self._chk(lib().hplx_pdfact_palg(
            self.h, ctypes.byref(palg), ctypes.c_void_p(P), lda, m, jb,
            ctypes.c_void_p(ipiv), row_offset))

    def pipid(self, jb: int, ipiv: int, permU: int) -> None:
        self._chk(lib().hplx_pipid(self.h, jb, ctypes.c_void_p(ipiv),
                                   ctypes.c_void_p(permU)))

    def dlaswp00N(self, jb: int, n: int, A: int, lda: int, permU: int) -> None:
        self._chk(lib().hplx_dlaswp00N(self.h, jb, n, ctypes.c_void_p(A), lda,
                                       ctypes.c_void_p(permU)))

    def pdupdate(self, A: int, lda: int, m: int, jb: int, n: int, permU: int,
                 W: int) -> None:
        self._chk(lib().hplx_pdupdate(self.h, ctypes.c_void_p(A), lda, m, jb,
                                      n, ctypes.c_void_p(permU),
                                      ctypes.c_void_p(W)))

    def dgemm(self, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc) -> None:
        self._chk(lib().hplx_dgemm(self.h, m, n, k, float(alpha),
                                   ctypes.c_void_p(A), lda, ctypes.c_void_p(B),
                                   ldb, float(beta), ctypes.c_void_p(C), ldc))

    def pdtrsv(self, n: int, nb: int, A: int, lda: int, x: int) -> None:
        self._chk(lib().hplx_pdtrsv(self.h, n, nb, ctypes.c_void_p(A), lda,
                                    ctypes.c_void_p(x)))

    def update_stats(self):
        ms, cnt, fl = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        self._chk(lib().hplx_update_stats(self.h, ctypes.byref(ms),
                                          ctypes.byref(cnt), ctypes.byref(fl)))
        return ms.value, cnt.value, fl.value

    def swap_timing(self):
        """(ms, launches, algorithmic bytes) of the last FP64 solve's row
        interchanges (hplx_swap_timing)."""
        ms, cnt, b = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        self._chk(lib().hplx_swap_timing(self.h, ctypes.byref(ms),
                                         ctypes.byref(cnt), ctypes.byref(b)))
        return ms.value, cnt.value, b.value

    def panel_timing(self):
        """(ms, panels, algorithmic bytes) of the last FP64 solve's panel
        factorisations (hplx_panel_timing)."""
        ms, cnt, b = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        self._chk(lib().hplx_panel_timing(self.h, ctypes.byref(ms),
                                          ctypes.byref(cnt), ctypes.byref(b)))
        return ms.value, cnt.value, b.value

    def update_timing(self, kind: int):
        """(ms, launches, flops) of one kind of update launch (hplx.h)."""
        ms, cnt, fl = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        self._chk(lib().hplx_update_timing(self.h, kind, ctypes.byref(ms),
                                           ctypes.byref(cnt), ctypes.byref(fl)))
        return ms.value, cnt.value, fl.value

    # ---- mixed precision (FP32/FP16 factors, BF16/FP8 update, FP64 IR) --------
    def mxp_alloc(self, n: int, nb: int,
                  storage: int = HPLX_STORE_FP32) -> MxpPMat:
        m = MxpPMat()
        self._chk(lib().hplx_mxp_alloc(self.h, n, nb, storage, ctypes.byref(m)))
        return m

    def mxp_free(self, mat: MxpPMat) -> None:
        self._chk(lib().hplx_mxp_free(self.h, ctypes.byref(mat)))

    def mxp_matgen(self, mat: MxpPMat, seed: int = HPL_ISEED,
                   diag_shift: float = 0.0) -> None:
        self._chk(lib().hplx_mxp_matgen(self.h, ctypes.byref(mat), seed,
                                        float(diag_shift)))

    def mxp_pdgetrf(self, palg: PAlg, opts: MxpOpts, mat: MxpPMat) -> int:
        return self._chk(lib().hplx_mxp_pdgetrf(
            self.h, ctypes.byref(palg), ctypes.byref(opts), ctypes.byref(mat)))

    def mxp_pdgetrs(self, mat: MxpPMat, x: int) -> None:
        self._chk(lib().hplx_mxp_pdgetrs(self.h, ctypes.byref(mat),
                                         ctypes.c_void_p(x)))

    def mxp_pdgesv(self, palg: PAlg, opts: MxpOpts, mat: MxpPMat):
        """Returns (rc, result dict): rc 0 when the refinement converged."""
        res = MxpResult()
        rc = self._chk(lib().hplx_mxp_pdgesv(
            self.h, ctypes.byref(palg), ctypes.byref(opts), ctypes.byref(mat),
            ctypes.byref(res)))
        return rc, res.as_dict()

    def mxp_residual(self, mat: MxpPMat, x: int) -> dict:
        out = (ctypes.c_double * 5)()
        self._chk(lib().hplx_mxp_residual(self.h, ctypes.byref(mat),
                                          ctypes.c_void_p(x), out))
        return dict(zip(["resid", "rnorm", "AnormI", "XnormI", "BnormI"],
                        list(out)))

    def mxp_update(self, update: int, terms: int, storage: int, m: int, n: int,
                   k: int, L: int, ldl: int, U: int, ldu: int, C: int,
                   ldc: int) -> None:
        self._chk(lib().hplx_mxp_update(self.h, update, terms, storage, m, n,
                                        k, ctypes.c_void_p(L), ldl,
                                        ctypes.c_void_p(U), ldu,
                                        ctypes.c_void_p(C), ldc))

    def spdfact(self, P: int, lda: int, m: int, jb: int, ipiv: int,
                nbmin: int = 32, nbdiv: int = 2, row_offset: int = 0) -> int:
        return self._chk(lib().hplx_spdfact(self.h, nbmin, nbdiv,
                                            ctypes.c_void_p(P), lda, m, jb,
                                            ctypes.c_void_p(ipiv), row_offset))

    # ---- whole-matrix helpers -------------------------------------------------
    def download_matrix(self, mat: PMat, rows: int | None = None,
                        cols: int | None = None) -> np.ndarray:
        rows = mat.n if rows is None else rows
        cols = mat.n + 1 if cols is None else cols
        out = np.empty((mat.ld, cols), dtype=np.float64, order="F")
        self._chk(lib().hplx_memcpy_d2h(self.h,
                                        out.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.c_void_p(mat.A), out.nbytes))
        return out[:rows, :]

    def download_vec(self, ptr: int, n: int, dtype=np.float64) -> np.ndarray:
        out = np.empty(n, dtype=dtype)
        if n:
            self._chk(lib().hplx_memcpy_d2h(
                self.h, out.ctypes.data_as(ctypes.c_void_p),
                ctypes.c_void_p(ptr), out.nbytes))
        return out

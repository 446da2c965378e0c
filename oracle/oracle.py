"""ctypes binding of the CPU oracle (oracle/hpl_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.  Each function
restates a piece of rocHPL's P=Q=1 HPL_pdgesv path; the reference file:line
each follows is cited in hpl_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhpl_oracle.so")
_lib = None

HPL_ISEED = 100  # include/hpl_ptest.hpp:57
EPS = 2.0 ** -53  # src/auxil/HPL_dlamch.cpp:106-108


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE,
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, i32, f64 = ctypes.c_uint64, ctypes.c_int, ctypes.c_double
        vp = ctypes.c_void_p
        L.hplo_jump.argtypes = [u64, u64, u64, ctypes.POINTER(u64),
                                ctypes.POINTER(u64)]
        L.hplo_start.argtypes = [u64]
        L.hplo_start.restype = u64
        L.hplo_state.argtypes = [u64, u64]
        L.hplo_state.restype = u64
        L.hplo_value.argtypes = [u64]
        L.hplo_value.restype = f64
        L.hplo_matgen_tile.argtypes = [i32, i32, i32, i32, i32, vp, i32, i32,
                                       f64]
        L.hplo_matgen.argtypes = [i32, i32, vp, i32, i32, f64]
        L.hplo_pdfact.argtypes = [vp, i32, i32, i32, i32, i32, vp]
        L.hplo_pdfact.restype = i32
        L.hplo_pdfact_v.argtypes = [vp, i32, i32, i32, i32, i32, i32, i32, vp]
        L.hplo_pdfact_v.restype = i32
        L.hplo_pipid.argtypes = [i32, vp, vp]
        L.hplo_laswp.argtypes = [vp, i32, i32, i32, vp]
        L.hplo_dlaswp00N.argtypes = [i32, i32, vp, i32, vp]
        L.hplo_trsv_upper.argtypes = [i32, vp, i32, vp]
        L.hplo_pdgesv.argtypes = [i32, i32, i32, i32, vp, i32, vp, vp]
        L.hplo_pdgesv.restype = i32
        L.hplo_pdgesv_v.argtypes = [i32, i32, i32, i32, i32, i32, vp, i32,
                                    vp, vp]
        L.hplo_pdgesv_v.restype = i32
        L.hplo_residual.argtypes = [i32, i32, f64, vp, vp]
        L.hplo_flops_ref.argtypes = [i32]
        L.hplo_flops_ref.restype = f64
        L.hplo_flops_hpl.argtypes = [i32]
        L.hplo_flops_hpl.restype = f64
        L.hplo_set_blas.argtypes = [vp, vp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def hpl_ld(n: int) -> int:
    """Leading dimension rule of src/matgen/HPL_pdmatgen.cpp:109-112."""
    return ((max(n, 0) + 95) // 128) * 128 + 32


def jump(k: int, mult: int = 6364136223846793005, iadd: int = 1):
    a, c = ctypes.c_uint64(), ctypes.c_uint64()
    lib().hplo_jump(k, mult, iadd, ctypes.byref(a), ctypes.byref(c))
    return a.value, c.value


def state(seed: int, k: int) -> int:
    return lib().hplo_state(seed, k)


def value(s: int) -> float:
    return lib().hplo_value(s)


def matgen(m: int, n: int, seed: int = HPL_ISEED, diag_shift: float = 0.0,
           lda: int | None = None) -> np.ndarray:
    """HPL_pdrandmat(GRID, m, n, NB, A, lda, seed) at P=Q=1, column-major
    (returned as a Fortran-ordered (lda, n) array sliced to m rows)."""
    lda = lda or max(m, 1)
    A = np.zeros((lda, n), dtype=np.float64, order="F")
    lib().hplo_matgen(m, n, _p(A), lda, seed, float(diag_shift))
    return A[:m, :]


def matgen_tile(M: int, i0: int, j0: int, m: int, n: int,
                seed: int = HPL_ISEED, diag_shift: float = 0.0) -> np.ndarray:
    A = np.zeros((max(m, 1), n), dtype=np.float64, order="F")
    lib().hplo_matgen_tile(M, i0, j0, m, n, _p(A), max(m, 1), seed,
                           float(diag_shift))
    return A[:m, :]


LEFT_LOOKING, CROUT, RIGHT_LOOKING = 301, 302, 303


def pdfact(P: np.ndarray, nbmin: int = 32, ndiv: int = 2,
           pfact: int = RIGHT_LOOKING, rfact: int = RIGHT_LOOKING):
    """Factor the panel P (m x jb, Fortran order) in place with the given
    panel (pfact) and recursion (rfact) variants; returns (ipiv
    panel-relative int32, info)."""
    assert P.flags.f_contiguous and P.dtype == np.float64
    m, jb = P.shape
    ipiv = np.zeros(max(jb, 1), dtype=np.int32)
    info = lib().hplo_pdfact_v(_p(P), max(m, 1), m, jb, nbmin, ndiv, pfact,
                               rfact, _p(ipiv))
    if info < 0:
        raise ValueError("unknown variant pfact=%d rfact=%d" % (pfact, rfact))
    return ipiv[:jb], info


def pipid(ipiv: np.ndarray) -> np.ndarray:
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    jb = ipiv.shape[0]
    permU = np.zeros(max(2 * jb, 1), dtype=np.int32)
    lib().hplo_pipid(jb, _p(ipiv), _p(permU))
    return permU[:2 * jb]


def laswp(A: np.ndarray, ipiv: np.ndarray) -> None:
    """Apply interchanges row i <-> ipiv[i] in order (in place)."""
    assert A.flags.f_contiguous and A.dtype == np.float64
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    lib().hplo_laswp(_p(A), max(A.shape[0], 1), A.shape[1], ipiv.shape[0],
                     _p(ipiv))


def dlaswp00N(A: np.ndarray, permU: np.ndarray) -> None:
    assert A.flags.f_contiguous and A.dtype == np.float64
    permU = np.ascontiguousarray(permU, dtype=np.int32)
    jb = permU.shape[0] // 2
    lib().hplo_dlaswp00N(jb, A.shape[1], _p(A), max(A.shape[0], 1), _p(permU))


def pdgesv(n: int, nb: int, nbmin: int = 32, ndiv: int = 2,
           seed: int = HPL_ISEED, diag_shift: float = 0.0,
           Ab: np.ndarray | None = None, pfact: int = RIGHT_LOOKING,
           rfact: int = RIGHT_LOOKING):
    """Full P=Q=1 HPL_pdgesv on the seeded [A|b].  Returns dict with the
    factored augmented matrix, global pivots, x and info."""
    lda = hpl_ld(n)
    if Ab is None:
        full = np.zeros((lda, n + 1), dtype=np.float64, order="F")
        lib().hplo_matgen(n, n + 1, _p(full), lda, seed, float(diag_shift))
    else:
        full = np.zeros((lda, n + 1), dtype=np.float64, order="F")
        full[:n, :] = Ab
    ipiv = np.zeros(max(n, 1), dtype=np.int32)
    x = np.zeros(max(n, 1), dtype=np.float64)
    info = lib().hplo_pdgesv_v(n, nb, nbmin, ndiv, pfact, rfact, _p(full),
                               lda, _p(ipiv), _p(x))
    if info < 0:
        raise ValueError("unknown variant pfact=%d rfact=%d" % (pfact, rfact))
    return {"LU": full[:n, :], "ipiv": ipiv[:n], "x": x[:n], "info": info,
            "lda": lda}


def residual(n: int, x: np.ndarray, seed: int = HPL_ISEED,
             diag_shift: float = 0.0) -> dict:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(9, dtype=np.float64)
    lib().hplo_residual(n, seed, float(diag_shift), _p(x), _p(out))
    keys = ["resid", "rnorm", "resid_r0_form", "AnormI", "Anorm1", "XnormI",
            "Xnorm1", "BnormI", "r0"]
    return dict(zip(keys, out.tolist()))


def flops_ref(n: int) -> float:
    return lib().hplo_flops_ref(n)


def flops_hpl(n: int) -> float:
    return lib().hplo_flops_hpl(n)


def use_host_blas(enable: bool = True) -> str:
    """Route the oracle's dgemm/dtrsm through scipy's host BLAS (used only for
    the timed CPU baseline).  Returns a description of the library."""
    if not enable:
        lib().hplo_set_blas(None, None)
        return "builtin loops"
    import scipy.linalg.cython_blas as cb
    ctypes.pythonapi.PyCapsule_GetPointer.restype = ctypes.c_void_p
    ctypes.pythonapi.PyCapsule_GetPointer.argtypes = [ctypes.py_object,
                                                      ctypes.c_char_p]
    ctypes.pythonapi.PyCapsule_GetName.restype = ctypes.c_char_p
    ctypes.pythonapi.PyCapsule_GetName.argtypes = [ctypes.py_object]

    def fp(name):
        cap = cb.__pyx_capi__[name]
        cname = ctypes.pythonapi.PyCapsule_GetName(cap)
        return ctypes.pythonapi.PyCapsule_GetPointer(cap, cname)

    lib().hplo_set_blas(fp("dgemm"), fp("dtrsm"))
    import scipy
    return "scipy %s bundled OpenBLAS (scipy.linalg.cython_blas)" % \
        scipy.__version__

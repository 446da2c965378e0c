"""The C-ABI library loads and exports every entry point include/hplx.h
declares; the Python mirror covers the same set.  CPU only (no device calls)."""
import ctypes
import os
import re
import subprocess

import pytest

import rochpl_amd as rp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hplx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(hplx_[a-z0-9_A-Z]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_expected_set():
    assert declared_functions() == sorted(rp.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(rp.LIB_PATH):
        pytest.skip("libhplx.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", rp.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (hplx_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_library_loads_and_binds_without_a_gpu():
    if not os.path.exists(rp.LIB_PATH):
        pytest.skip("libhplx.so not built")
    L = rp.lib()
    for name in rp.EXPORTS:
        assert hasattr(L, name)
    p = rp.palg_default()
    # reference defaults, src/HPL_pdinfo.cpp:487-562
    assert (p.pfact, p.rfact, p.nbmin, p.nbdiv, p.depth) == \
        (rp.HPLX_RIGHT_LOOKING, rp.HPLX_RIGHT_LOOKING, 32, 2, 1)
    assert (p.L1notran, p.Unotran, p.equil) == (0, 0, 0)


def test_struct_layouts_match_header():
    # hplx_palg: 12 ints, a double at offset 48, an int, padded -> 64 bytes
    assert ctypes.sizeof(rp.PAlg) == 64
    assert rp.PAlg.frac.offset == 48
    # hplx_pmat: 3 pointers, 6 ints, a pointer -> 56 bytes on LP64
    assert ctypes.sizeof(rp.PMat) == 56
    assert rp.PMat.ws.offset == 48


def test_struct_layouts_match_a_c_compiler(tmp_path):
    """sizeof / offsetof of every header struct, as gcc lays them out, equal
    the ctypes mirrors'."""
    import shutil
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    structs = {"hplx_palg": rp.PAlg, "hplx_pmat": rp.PMat,
               "hplx_mxp_opts": rp.MxpOpts, "hplx_mxp_pmat": rp.MxpPMat,
               "hplx_mxp_result": rp.MxpResult}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hplx.h"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));'
                         % (cname, fname, cname, fname))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, \
                (cname, fname)


def test_errors_without_device_are_codes_not_exits():
    if not os.path.exists(rp.LIB_PATH):
        pytest.skip("libhplx.so not built")
    h = ctypes.c_void_p()
    rc = rp.lib().hplx_ctx_create(10 ** 6, ctypes.byref(h))
    assert rc < 0 and not h.value
    assert rp.lib().hplx_last_error(None) == b"null context"


@pytest.mark.parametrize("field,value", [
    ("depth", 0), ("depth", 2),                      # src/HPL_pdinfo.cpp:930-935
    ("fswap", rp.HPLX_SWAP00), ("fswap", rp.HPLX_SW_MIX),  # :948-955
    ("Unotran", 1),                                  # :979-986
    ("equil", 1),                                    # :996-1003
])
def test_palg_settings_the_reference_refuses_are_unsupported(field, value):
    """HPL_pdinfo refuses these HPL.dat settings; hplx_palg_check (the check
    hplx_pdgesv / hplx_pdgetrf run first) reports them without a device."""
    if not os.path.exists(rp.LIB_PATH):
        pytest.skip("libhplx.so not built")
    p = rp.palg_default()
    assert rp.lib().hplx_palg_check(ctypes.byref(p)) == rp.HPLX_SUCCESS
    setattr(p, field, value)
    assert rp.lib().hplx_palg_check(ctypes.byref(p)) == rp.HPLX_ERR_UNSUPPORTED


@pytest.mark.parametrize("field,value", [
    ("pfact", rp.HPLX_LEFT_LOOKING), ("pfact", rp.HPLX_CROUT),
    ("rfact", rp.HPLX_LEFT_LOOKING), ("rfact", rp.HPLX_CROUT),
    ("L1notran", 1)])
def test_palg_panel_variants_are_accepted(field, value):
    """Every PFACT / RFACT of HPL.dat and L1 in either form
    (src/HPL_pddriver.cpp:164-200) are built."""
    if not os.path.exists(rp.LIB_PATH):
        pytest.skip("libhplx.so not built")
    p = rp.palg_default()
    setattr(p, field, value)
    assert rp.lib().hplx_palg_check(ctypes.byref(p)) == rp.HPLX_SUCCESS


def test_palg_invalid_recursion_and_null():
    if not os.path.exists(rp.LIB_PATH):
        pytest.skip("libhplx.so not built")
    for field, value in (("nbmin", 0), ("nbdiv", 1), ("pfact", 300),
                         ("rfact", 304), ("L1notran", 2)):
        p = rp.palg_default()
        setattr(p, field, value)
        assert rp.lib().hplx_palg_check(ctypes.byref(p)) == rp.HPLX_ERR_INVALID
    assert rp.lib().hplx_palg_check(None) == rp.HPLX_ERR_INVALID
    # frac is not read at P=1 (src/panel/HPL_pdpanel_init.cpp:176-184)
    p = rp.palg_default()
    p.frac = 0.6
    assert rp.lib().hplx_palg_check(ctypes.byref(p)) == rp.HPLX_SUCCESS

"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
the xhplx driver's HPL.dat parser, echo and error paths (`--parse-only`, so
no GPU call is made), and every entry point of the C oracle.  Each binary
is built by its own Makefile's `sanitize` target; a sanitizer report aborts
the run with a non-zero status."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "rochpl_amd", "csrc")
LIB = os.path.join(ROOT, "rochpl_amd", "libhplx.so")
XSAN = os.path.join(CSRC, "_build", "xhplx_san")
OSAN = os.path.join(ROOT, "oracle", "_san", "oracle_san")
GOLDEN = os.path.join(ROOT, "tests", "golden")
# the HIP runtime the driver links keeps allocations until exit: leaks are
# not this check's subject
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _make(target_dir):
    r = subprocess.run(["make", "-s", "-C", target_dir, "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr


def _clean(r):
    text = r.stdout + r.stderr
    assert "AddressSanitizer" not in text and "runtime error" not in text, text
    return text


def test_oracle_under_sanitizers():
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    _make(os.path.join(ROOT, "oracle"))
    r = subprocess.run([OSAN], capture_output=True, text=True, timeout=300,
                       env=ENV)
    _clean(r)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout


def _edit(tmp_path, lines_from, idx, value):
    lines = open(lines_from).read().splitlines()
    lines[idx] = value
    p = tmp_path / "HPL.dat"
    p.write_text("\n".join(lines) + "\n")
    return str(p)


def test_driver_parser_under_sanitizers(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("libhplx.so not built")
    _make(CSRC)
    cfg1 = os.path.join(GOLDEN, "HPL_config1.dat")
    multi = os.path.join(GOLDEN, "HPL_multi.dat")
    ok_runs = [["-i", cfg1, "--parse-only"], ["-i", multi, "--parse-only"],
               ["-N", "2048", "-NB", "128", "--parse-only"], ["-h"],
               ["--version"]]
    for args in ok_runs:
        r = subprocess.run([XSAN] + args, cwd=tmp_path, capture_output=True,
                           text=True, timeout=120, env=ENV)
        _clean(r)
        assert r.returncode == 0, (args, r.stdout + r.stderr)
    # error paths: illegal values, a short file, a missing file, bad flags
    bad = [
        ["-i", _edit(tmp_path, cfg1, 25, "0            DEPTHs (>=0)"),
         "--parse-only"],
        ["-i", _edit(tmp_path, cfg1, 5, "21           # of problems sizes (N)"),
         "--parse-only"],
        ["-i", _edit(tmp_path, cfg1, 17, "0            NBMINs"),
         "--parse-only"],
        ["-i", str(tmp_path / "missing.dat"), "--parse-only"],
        ["-N", "0", "--parse-only"], ["-P", "2", "--parse-only"],
        ["--no-such-flag"],
    ]
    short = tmp_path / "short.dat"
    short.write_text("\n".join(open(cfg1).read().splitlines()[:9]) + "\n")
    bad.append(["-i", str(short), "--parse-only"])
    for args in bad:
        r = subprocess.run([XSAN] + args, cwd=tmp_path, capture_output=True,
                           text=True, timeout=120, env=ENV)
        _clean(r)
        assert r.returncode != 0, args

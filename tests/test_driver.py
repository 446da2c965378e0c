"""xhplx driver surface (HPL.dat parser, command line, error behaviour),
exercised with --parse-only so no GPU is touched.  CPU only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
XHPLX = os.path.join(ROOT, "rochpl_amd", "xhplx")
CONFIG1 = os.path.join(ROOT, "tests", "golden", "HPL_config1.dat")

pytestmark = pytest.mark.skipif(not os.path.exists(XHPLX),
                                reason="xhplx not built")


def run(args, cwd):
    return subprocess.run([XHPLX, "--parse-only"] + args, cwd=cwd,
                          capture_output=True, text=True, timeout=60)


def edit(tmp_path, line_no, value):
    lines = open(CONFIG1).read().splitlines()
    lines[line_no - 1] = value
    p = tmp_path / "HPL.dat"
    p.write_text("\n".join(lines) + "\n")
    return str(p)


def test_config1_echo(tmp_path):
    r = run(["-i", CONFIG1], tmp_path)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "N      :    4096" in out
    assert "NB     :     256" in out
    assert "PFACT  :   Right" in out and "RFACT  :   Right" in out
    assert "NBMIN  :      32" in out and "NDIV   :       2" in out
    assert "DEPTH  :       1" in out
    assert "than           16.0" in out


def test_multi_value_lines(tmp_path):
    p = tmp_path / "HPL.dat"
    lines = open(CONFIG1).read().splitlines()
    lines[4] = "3            # of problems sizes (N)"
    lines[5] = "1000 2000 3000  Ns"
    lines[6] = "2            # of NBs"
    lines[7] = "128 256      NBs"
    p.write_text("\n".join(lines) + "\n")
    r = run(["-i", str(p)], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "N      :    1000     2000     3000" in r.stdout
    assert "NB     :     128      256" in r.stdout


@pytest.mark.parametrize("line,value,msg", [
    (25, "0            DEPTHs (>=0)", "Lookahead depth must be 1"),
    (26, "0            SWAP", "Only long swapping"),
    (29, "1            U  in", "U  in no-transposed form unsupported"),
    (30, "1            Equilibration", "Equilibration currently unsupported"),
    (17, "0            NBMINs", "Value of NBMIN less than 1"),
    (19, "1            NDIVs", "Value of NDIV less than 2"),
    (5, "21           # of problems sizes (N)", "greater than 20"),
])
def test_illegal_input_exits_1(tmp_path, line, value, msg):
    r = run(["-i", edit(tmp_path, line, value)], tmp_path)
    assert r.returncode == 1
    assert msg in r.stderr
    assert "Illegal input in file" in r.stderr


def test_missing_file(tmp_path):
    r = run(["-i", str(tmp_path / "nope.dat")], tmp_path)
    assert r.returncode == 1 and "cannot open file" in r.stderr


def test_command_line_mode_writes_hpl_out(tmp_path):
    r = run(["-N", "2048", "-NB", "128"], tmp_path)
    assert r.returncode == 0, r.stderr
    text = (tmp_path / "HPL.out").read_text()
    assert "N      :    2048" in text and "NB     :     128" in text


def test_grid_larger_than_one_gpu_is_refused(tmp_path):
    r = run(["-P", "2", "-Q", "1", "-N", "1000"], tmp_path)
    assert r.returncode == 1 and "P=Q=1" in r.stderr


def test_illegal_command_line_values(tmp_path):
    for args in (["-N", "0"], ["-NB", "0"], ["-P", "0"]):
        r = run(args, tmp_path)
        assert r.returncode == 1 and "Illegal value" in r.stderr


def test_help_and_version(tmp_path):
    h = subprocess.run([XHPLX, "-h"], capture_output=True, text=True)
    assert h.returncode == 0 and "-NB" in h.stdout
    v = subprocess.run([XHPLX, "--version"], capture_output=True, text=True)
    assert v.returncode == 0 and "version" in v.stdout


def test_echo_wraps_long_lists_like_the_reference(tmp_path):
    """HPL_pdinfo's echo (src/HPL_pdinfo.cpp:1288-1310): "%8d " per value,
    eight to a line, continuation lines indented by eight blanks."""
    p = tmp_path / "HPL.dat"
    lines = open(CONFIG1).read().splitlines()
    ns = list(range(1000, 11000, 1000))        # 10 values
    lines[4] = "10           # of problems sizes (N)"
    lines[5] = " ".join(str(v) for v in ns) + "  Ns"
    p.write_text("\n".join(lines) + "\n")
    r = run(["-i", str(p)], tmp_path)
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    i = out.index("N      :" + "".join("%8d " % v for v in ns[:8]))
    assert out[i + 1] == "        " + "".join("%8d " % v for v in ns[8:])
    assert "NB     :" + "%8d " % 256 in out

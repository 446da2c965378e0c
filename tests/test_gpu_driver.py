"""End-to-end run of the xhplx driver on the config-1 HPL.dat (N=4096,
NB=256, P=Q=1): the reference's own acceptance check is the PASSED line."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
XHPLX = os.path.join(ROOT, "rochpl_amd", "xhplx")
CONFIG1 = os.path.join(ROOT, "tests", "golden", "HPL_config1.dat")

pytestmark = pytest.mark.gpu


def test_xhplx_config1_passes(tmp_path):
    print("[gpu-case] xhplx -i HPL_config1.dat", flush=True)
    r = subprocess.run([XHPLX, "-i", CONFIG1], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    w = re.search(r"^WC1[0-5]R2R32\s+4096\s+256\s+1\s+1\s+(\S+)\s+(\S+)$", out,
                  re.M)
    assert w, out
    assert float(w.group(1)) > 0 and float(w.group(2)) > 0
    m = re.search(r"\)\*N\)=\s+(\S+) \.\.\.\.\.\. PASSED", out)
    assert m and float(m.group(1)) < 16.0, out
    assert "1 tests completed and passed residual checks" in out

"""The packed CSR entries of the sweeps store a ray count per entry and
recompute the weight count / den with a reciprocal and one FMA correction
step.  The kernels rely on that step giving the correctly rounded quotient
(the value of the FP32 weight array) for every 0 <= count <= den <= 65535;
this checks it exhaustively with the host's IEEE FMA."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_reciprocal_correction_is_correctly_rounded(tmp_path):
    exe = tmp_path / "packed_weight_check"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(HERE / "packed_weight_check.c"), "-lm"],
                   check=True)
    r = subprocess.run([str(exe), "65535"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert "mismatches=0" in r.stdout

"""The OpenFOAM adapter (hipLduSolver, hipLduPreconditioner) compiles against
the OpenFOAM headers: check_compile.sh syntax-checks both files with g++ as
wmake would, through per-library lnInclude directories (OpenFOAM,
OSspecific/POSIX, finiteVolume, meshTools and the libraries their headers
include).  Needs an OpenFOAM source tree; skipped where there is none."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "openfoam-dev_amd", "adapter", "check_compile.sh")
FOAM_SRC = os.environ.get("FOAM_SRC", "/root/reference/src")


@pytest.mark.skipif(not os.path.isdir(os.path.join(FOAM_SRC, "OpenFOAM")),
                    reason="no OpenFOAM source tree to compile the adapter against")
def test_adapter_compiles_against_openfoam_headers():
    out = subprocess.run(["sh", SCRIPT, FOAM_SRC], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-4000:]
    assert "hipLduSolver.C: OK" in out.stdout
    assert "hipLduPreconditioner.C: OK" in out.stdout

"""The C++ operator surface (include/frank_amd.hpp) driven like the reference's
drivers: BLR-LU solve, dispatch errors, dense factorizations, rounded addition
and tiled BLR-QR (tests/cxx/surface_test.cpp, built by frank_amd/build.py)."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent.parent / "frank_amd" / "bin" / "surface_test"


def test_cxx_surface_driver(torch_gpu):
    assert EXE.exists(), "build the library first (frank_amd/build.py)"
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    print(res.stderr)
    assert res.returncode == 0
    lines = [l for l in res.stdout.splitlines() if l.startswith(("ok", "FAIL"))]
    assert len(lines) == 12 and all(l.startswith("ok") for l in lines)

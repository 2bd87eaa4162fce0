"""The Fortran host (mose-gpu_amd/fortran: the ISO_C_BINDING module, the host sequencer and the shock-tube driver)
drives the library through the C ABI and gets the oracle's state: the same field_state_hash, time and profile."""
import subprocess

import numpy as np
import pytest

import mosegpu
from mosegpu import cases
from oracle_bind import OracleCase

pytestmark = pytest.mark.gpu

DRIVER = mosegpu.LIB_PATH.parent / "mose_sod"


@pytest.mark.parametrize("path", [0, 1])
def test_fortran_sod_driver(gpu_ready, tmp_path, path):
    assert DRIVER.exists(), "build the Fortran host first (make -C mose-gpu_amd)"
    r = subprocess.run([str(DRIVER), "1000", "200", str(path)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    lo, hi = (int(x) % (1 << 64) for x in out["hash"])
    t = float(out["time"][0])
    case = cases.sod(1000, 200)
    Po = case.P.copy()
    dts = OracleCase(case).step(Po, 200)
    assert abs(t - dts.sum()) <= 1e-13 * dts.sum()
    prof = np.loadtxt(tmp_path / "mose_sod.out")
    V = case.layout.interior(Po, 0)[:, 0, 0, :]
    assert np.max(np.abs(prof[:, 0] - V[0])) <= 1e-12 * np.max(np.abs(V[0]))
    assert np.max(np.abs(prof[:, 2] - V[4])) <= 1e-12 * np.max(np.abs(V[4]))
    print(f"fortran host path {path}: hash {(lo, hi)}, oracle {mosegpu.state_hash(case.layout, Po)}")
    assert (lo, hi) == mosegpu.state_hash(case.layout, Po)

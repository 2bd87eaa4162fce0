"""The Fortran host (mose-gpu_amd/fortran: the ISO_C_BINDING module, the host sequencer, the shock-tube and chemistry
drivers) drives the library through the C ABI and gets the oracle's state: the same field_state_hash, time and
profile, and the same chemistry state bit for bit."""
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


def write_mechanism(path, m):
    with open(path, "wb") as f:
        np.array([m.ns, m.n_arrh, m.n_troe, m.n_lind, 0, m.cp.shape[1] - 1], dtype=np.int32).tofile(f)
        for a in (m.wm, m.cp, m.h, m.ni1, m.ni2, m.eps, m.kf, m.kb, m.troe, m.lind):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)


def test_fortran_chemistry_driver(gpu_ready, tmp_path):
    """mose_chem_run (MOSE_GPU_Chemistry_Newstate through the ISO_C_BINDING module) on 96 hydrogen-air cells: the
    oracle's RODAS state and hr, bit for bit."""
    from mosegpu import chem
    from oracle_bind import OracleChem
    drv = mosegpu.LIB_PATH.parent / "mose_chem_run"
    assert drv.exists(), "build the Fortran host first (make -C mose-gpu_amd)"
    mech = chem.h2_air_mechanism()
    st = chem.h2_air_cells(96, seed=41)
    write_mechanism(tmp_path / "mech.bin", mech)
    with open(tmp_path / "cells.bin", "wb") as f:
        np.array([st.shape[1]], dtype=np.int64).tofile(f)
        np.ascontiguousarray(st).tofile(f)          # [ns+1][ncell] = Fortran state(ncell, ns+1)
    r = subprocess.run([str(drv), "mech.bin", "cells.bin", "out.bin", "1e-6", "1", "1e-8", "1e-12"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    n = st.shape[1]
    dev, hr = raw[:10 * n].reshape(10, n), raw[10 * n:]
    ref, hr_o, idid_o, _ = OracleChem(mech, nthreads=16).advance(st, 1e-6)
    assert np.all(idid_o == 1)
    print(f"fortran chemistry driver: {r.stdout.strip()}; max |dev - oracle| {np.abs(dev - ref).max():.3e}")
    assert np.array_equal(dev, ref)
    assert np.array_equal(hr, hr_o)


def run_case_driver(tmp_path, case, nsteps, path=0, tile=(0, 0, 0), mech=None, opts=None):
    from mosegpu import caseio
    drv = mosegpu.LIB_PATH.parent / "mose_case_run"
    assert drv.exists(), "build the Fortran host first (make -C mose-gpu_amd)"
    caseio.write_fortran_case(case, tmp_path / "case.bin", nsteps, path, tile, mech, opts)
    r = subprocess.run([str(drv), "case.bin", "out.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return caseio.read_fortran_result(case, tmp_path / "out.bin") + (r.stdout.strip(),)


@pytest.mark.parametrize("path", [0, 1])
def test_fortran_case_driver_taylor_green(gpu_ready, tmp_path, path):
    """mose_case_run (the whole residual path through the ISO_C_BINDING module from a case file) on the viscous
    Taylor-Green vortex 16^3: the oracle's state and time."""
    from test_gpu_walls_sa import compare
    case = cases.taylor_green(16, 4)
    P, t, steps, out = run_case_driver(tmp_path, case, 4, path)
    Po = case.P.copy()
    dts = OracleCase(case).step(Po, 4)
    assert steps == 4 and abs(t - dts.sum()) <= 1e-13 * dts.sum()
    assert compare(case, P, Po, f"fortran case driver, TGV 16^3, path {path} ({out})") <= 1e-12


def test_fortran_case_driver_reacting(gpu_ready, tmp_path):
    """mose_case_run on the nine-species reacting box: MOSE_GPU_Chemistry_Attach couples the mechanism, the steps are
    Strang-split, the state is the oracle's to the chemistry bar."""
    from mosegpu import chem
    from oracle_bind import OracleChem
    from test_gpu_walls_sa import compare
    mech = chem.h2_air_mechanism()
    case = cases.reacting_box((12, 10, 8), T_range=(1200.0, 1700.0))
    P, t, steps, out = run_case_driver(tmp_path, case, 3, 0, (0, 0, 0), mech, chem.ode_opts())
    o = OracleCase(case)
    o.set_chemistry(OracleChem(mech, nthreads=16))
    Po = case.P.copy()
    dts = o.step(Po, 3)
    assert o.chem_failed() == 0
    assert steps == 3 and abs(t - dts.sum()) <= 1e-13 * dts.sum()
    assert compare(case, P, Po, f"fortran case driver, reacting box 12x10x8 x 3 Strang steps ({out})") <= 1e-8


@pytest.mark.parametrize("name,nsteps,path", [("isentropic-vortex", 20, 0), ("isentropic-vortex", 20, 1),
                                              ("shu-osher", 330, 0), ("diff1", 100, 0), ("noh87", 100, 0),
                                              ("sod79", 100, 0), ("forward-step", 100, 0)])
def test_fortran_case_dir_runs_the_reference_case(gpu_ready, tmp_path, name, nsteps, path):
    """mose_case_dir reads the reference's own case directory (input.ini and INPUT/) in Fortran, builds the geometry,
    steps the case through MOSE_GPU_Explicit_Step and writes the state: the oracle's state and time on the same
    inputs (read by the Python loader)."""
    import os
    from fortran_case_dir import GOLD, expand, read_result, run
    from mosegpu import caseio
    from test_gpu_walls_sa import compare
    d = expand(name, tmp_path)
    r = run(d, nsteps, str(tmp_path / "out.bin"), path)
    assert r.returncode == 0, r.stdout + r.stderr
    case = caseio.load_case(os.path.join(GOLD, name), nsteps)
    P, t, steps = read_result(str(tmp_path / "out.bin"), case.layout.NC, case.nprim)
    Po = case.P.copy()
    dts = OracleCase(case, nthreads=16).step(Po, nsteps)
    assert steps == nsteps and abs(t - dts.sum()) <= 1e-13 * dts.sum()
    assert compare(case, P, Po, f"fortran case directory {name}, path {path} x {nsteps} ({r.stdout.strip()})") <= 1e-12


def test_fortran_case_dir_runs_to_the_time_threshold(gpu_ready, tmp_path):
    """Noh87 has no iter-threshold: mose_case_dir with nsteps -1 steps while the time is below its time-threshold
    (the reference's time loop), and the oracle stepped the same way gives the same state, time and step count."""
    import os
    from fortran_case_dir import GOLD, expand, read_result, run
    from mosegpu import caseio
    from test_gpu_walls_sa import compare
    d = expand("noh87", tmp_path)
    r = run(d, -1, str(tmp_path / "out.bin"))
    assert r.returncode == 0, r.stdout + r.stderr
    case = caseio.load_case(os.path.join(GOLD, "noh87"))
    tmax = case.meta["time_threshold"]
    assert case.nsteps == 0 and tmax > 0
    P, t, steps = read_result(str(tmp_path / "out.bin"), case.layout.NC, case.nprim)
    Po = case.P.copy()
    o = OracleCase(case, nthreads=16)
    time, n = 0.0, 0
    while time < tmax:
        time += float(o.step(Po, 1).sum())
        n += 1
    print(f"noh87 to t = {tmax}: {steps} steps on the device, {n} in the oracle", flush=True)
    assert steps == n and abs(t - time) <= 1e-13 * time
    assert compare(case, P, Po, f"fortran case directory noh87 to the time threshold ({r.stdout.strip()})") <= 1e-12

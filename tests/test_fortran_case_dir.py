"""The reference's case format read by the Fortran host (mose-gpu_amd/fortran/mose_case_files.f90, program
mose_case_dir): input.ini, phase.txt, thermo.dat / transport.dat, ic.tec and bc.txt of the reference's own test
cases, turned into the C ABI's arrays.  The program's dump mode (no GPU) writes what it read: the scheme, the tables,
the geometry from mose_metrics_build, the sorted records and the initial arena must equal the Python loader's
(caseio.load_case) bit for bit.  Inputs off the path stop the program with a message that names them."""
import os

import numpy as np
import pytest

from fortran_case_dir import CASES, GOLD, expand, read_dump, run
from mosegpu import caseio


@pytest.mark.parametrize("name", CASES)
def test_fortran_reads_the_reference_case(tmp_path, name):
    d = expand(name, tmp_path)
    r = run(d, 5, str(tmp_path / "dump.bin"), "dump")
    assert r.returncode == 0, r.stdout + r.stderr
    case = caseio.load_case(os.path.join(GOLD, name), 5)
    got = read_dump(str(tmp_path / "dump.bin"), bool(case.scheme.viscous), case.nprim)
    assert got["scheme"] == bytes(case.scheme)
    assert got["ndir"] == case.meta["ndir"]
    assert np.array_equal(got["dims"], case.layout.dims_arr)
    t = case.tables
    for k, ref in (("wm", t.wm), ("cp", t.cp), ("h", t.h)) + ((("mu", t.mu), ("k", t.k)) if case.scheme.viscous else ()):
        assert np.array_equal(got[k], np.ravel(ref)), k
    g = case.geom
    for k in ("dl", "m", "vol", "fm1", "fm2", "fm3"):
        assert np.array_equal(got[k], getattr(g, k)), k
    assert got["rec"].tobytes() == np.ascontiguousarray(case.records).tobytes()
    assert got["ff"].tobytes() == np.ascontiguousarray(case.farfield).tobytes()
    assert np.array_equal(got["P"], case.P.ravel())
    print(f"{name}: {case.layout.dims} blocks, {len(case.records)} records, nprim {case.nprim}: equal", flush=True)


def _edit(path, old, new):
    s = open(path).read()
    assert old in s
    open(path, "w").write(s.replace(old, new, 1))


@pytest.mark.parametrize("what,file,old,new,msg", [
    ("boundary type", "INPUT/bc.txt", "201", "501", "boundary type 501"),
    ("local time stepping", "input.ini", "time-accurate = true", "time-accurate = false", "local time stepping"),
    ("solver", "input.ini", "riemann-solver = HLLC", "riemann-solver = AUSM+M", "riemann-solver = AUSM+M"),
    ("preconditioned update", "input.ini", "integration-variables = cons", "integration-variables = prec",
     "integration-variables = prec"),
])
def test_fortran_refuses_what_is_off_the_path(tmp_path, what, file, old, new, msg):
    d = expand("isentropic-vortex", tmp_path)
    _edit(os.path.join(d, file), old, new)
    r = run(d, 5, str(tmp_path / "dump.bin"), "dump")
    assert r.returncode != 0, what
    assert msg in r.stderr, r.stderr


def test_fortran_reads_the_primitive_update_key(tmp_path):
    """integration-variables = prim: the Fortran reader's scheme equals the Python loader's (integration = 1)."""
    d = expand("isentropic-vortex", tmp_path)
    _edit(os.path.join(d, "input.ini"), "integration-variables = cons", "integration-variables = prim")
    r = run(d, 5, str(tmp_path / "dump.bin"), "dump")
    assert r.returncode == 0, r.stdout + r.stderr
    case = caseio.load_case(d, 5)
    assert case.scheme.integration == 1
    got = read_dump(str(tmp_path / "dump.bin"), bool(case.scheme.viscous), case.nprim)
    assert got["scheme"] == bytes(case.scheme)

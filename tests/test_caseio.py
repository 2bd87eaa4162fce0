"""The reference's own test-case inputs (tests/golden/cases, copied from test/2D/euler/isentropic-vortex by
tests/golden/cases/make_cases.py) read by mosegpu.caseio, and the oracle run on the case that results.

Pins: what the files state (64 x 64 x 1 cells, periodic MOSE 201 faces in x and y, MOSE 0 faces in the one-cell
direction, HLLC + van Leer, SSP-RK3 at CFL 0.5, the constant-cp air table); exact conservation of mass and momentum
on the periodic box (every face flux enters one cell and leaves another); the vortex drifting with the mean flow
(1, 1); and the refusals of inputs this path does not run.
"""
import os
import shutil

import numpy as np
import pytest

from mosegpu import BC_CONNECTION, BC_NONE, caseio
from mosegpu.cases import CP_AIR, R_AIR
from oracle_bind import OracleCase

HERE = os.path.dirname(os.path.abspath(__file__))
VORTEX = os.path.join(HERE, "golden", "cases", "isentropic-vortex")


@pytest.fixture(scope="module")
def vortex():
    return caseio.load_case(VORTEX, nsteps=20)


def test_vortex_inputs(vortex):
    c = vortex
    assert c.layout.dims == [(64, 64, 1)]
    assert c.meta["ndir"] == 2
    typ = c.records["typ"]
    assert (typ == BC_CONNECTION).sum() == 4 * 64 and (typ == BC_NONE).sum() == 2 * 64 * 64
    per = c.records[typ == BC_CONNECTION]
    assert np.all(per["periodic"] == 1) and np.all(per["bs"] == 0)
    # the donor of a face-1 cell is the face-2 cell of the same row, and the other way round
    f1 = per[per["f"] == 1]
    assert np.all(f1["fs"] == 2) and np.all(f1["is_"] == 64) and np.all(f1["js"] == f1["j"])
    f3 = per[per["f"] == 3]
    assert np.all(f3["fs"] == 4) and np.all(f3["js"] == 64) and np.all(f3["is_"] == f3["i"])
    # records grouped by block, then face
    assert np.all(np.diff(c.records["f"]) >= 0)
    s = c.scheme
    assert (s.limiter, s.riemann, s.n_rk, s.viscous, s.nrans, s.cfl) == (2, 1, 3, 0, 0, 0.5)
    # thermo.dat: constant cp of air from T = 1 K, FLINT's row 0 = cp(1), h = 0
    assert np.all(c.tables.cp[0] == CP_AIR) and c.tables.h[0, 0] == 0.0 and c.tables.h[0, 2] == 2 * CP_AIR
    assert c.tables.wm[0] == 28.970418
    # the box: 10 x 10 in x and y, one cell of 1e-3 in z
    nd = c.nodes[0]
    assert nd.shape == (2, 65, 65, 3)
    assert nd[0, 0, 0, 0] == 0.0 and nd[0, 0, 64, 0] == 10.0 and abs(nd[1, 0, 0, 2] - nd[0, 0, 0, 2] - 1e-3) < 1e-9
    V = c.layout.interior(c.P, 0)
    assert V.shape == (5, 1, 64, 64)
    assert np.all(V[3] == 0.0) and V[0].max() == 1.0 and V[4].max() == 1.0


def test_vortex_oracle_conserves_and_advects(vortex):
    c = vortex
    vol = c.layout.interior(c.geom.vol[None, :], 0)[0]

    def totals(P):
        V = c.layout.interior(P, 0)
        return np.array([np.sum(V[0] * vol), np.sum(V[0] * V[1] * vol), np.sum(V[0] * V[2] * vol)])

    def centre(P):
        # centroid of the density deficit, unwrapped about the initial centre (5, 5)
        V = c.layout.interior(P, 0)[0, 0]
        xc = (np.arange(64) + 0.5) * (10.0 / 64)
        wgt = (1.0 - V).clip(min=0.0)
        X, Y = np.meshgrid(xc, xc, indexing="xy")
        return np.array([np.sum(wgt * X) / wgt.sum(), np.sum(wgt * Y) / wgt.sum()])

    P = c.P.copy()
    t0, c0 = totals(P), centre(P)
    dts = OracleCase(c, nthreads=8).step(P, 20)
    t1, c1 = totals(P), centre(P)
    assert np.all(np.abs(t1 - t0) <= 1e-13 * np.abs(t0))
    T = dts.sum()
    assert np.all(np.abs((c1 - c0) - T) < 0.02 * 10.0 / 64 + 0.05 * T)      # moves with the mean flow (1, 1)
    assert np.all(c.layout.interior(P, 0)[3] == 0.0)                          # no flux through the MOSE 0 faces
    # the case is cold in these units: T = p / (rho R) < 1 K, the table's row 0 (as the reference reads it)
    V = c.layout.interior(P, 0)
    assert np.max(V[4] / (V[0] * R_AIR)) < 1.0


@pytest.mark.parametrize("key,value,msg", [
    ("time-accurate", "false", "local time stepping"),
    ("riemann-solver", "AUSM+M", "AUSM"),
    ("integration-variables", "prec", "integration-variables"),
])
def test_refusals(tmp_path, key, value, msg):
    d = tmp_path / "case"
    shutil.copytree(VORTEX, d)
    ini = (d / "input.ini").read_text().splitlines()
    out = [(f"{key} = {value}" if ln.split("=")[0].strip() == key else ln) for ln in ini]
    if not any(ln.split("=")[0].strip() == key for ln in ini):
        out = [ln + (f"\n{key} = {value}" if ln.strip() == "[MOSE-Numerics]" else "") for ln in out]
    (d / "input.ini").write_text("\n".join(out) + "\n")
    with pytest.raises(caseio.CaseFormatError, match=msg):
        caseio.load_case(str(d))


def test_primitive_update_key(tmp_path):
    """integration-variables = prim selects the primitive update (Assign_Integration_Variables, Lib_Newstate.f90:43-60)."""
    d = tmp_path / "case"
    shutil.copytree(VORTEX, d)
    ini = (d / "input.ini").read_text()
    assert "integration-variables = cons" in ini
    (d / "input.ini").write_text(ini.replace("integration-variables = cons", "integration-variables = prim"))
    assert caseio.load_case(str(d)).scheme.integration == 1
    assert caseio.load_case(str(VORTEX)).scheme.integration == 0


def test_unsupported_boundary_type(tmp_path):
    d = tmp_path / "bc.txt"
    d.write_text("       1       1       1       1       3     303\n")
    with pytest.raises(caseio.CaseFormatError, match="303"):
        caseio.read_bc_open(str(d))


def test_symmetry_type(tmp_path):
    """MOSE 300 (Euler symmetry, no property line) is the internal type 3 (IO_BC_Open.f90:191-193)."""
    d = tmp_path / "bc.txt"
    d.write_text("       1       1       1       1       3     300\n       1       2       1       1       3     300\n")
    r = caseio.read_bc_open(str(d))
    assert list(r["typ"]) == [3, 3] and list(r["i"]) == [1, 2]


def test_tecplot_block_reader(tmp_path):
    """A two-zone BLOCK file with nodal and cell-centred variables, several values per line."""
    p = tmp_path / "ic.tec"
    z1 = np.arange(3 * 2 * 2, dtype=float)
    lines = ['VARIABLES = "x" "y" "z" "q"',
             'ZONE T="a", I=3, J=2, K=2, DATAPACKING=BLOCK, VARLOCATION=([1-3]=NODAL,[4]=CELLCENTERED)']
    for v in (z1, z1 + 100, z1 + 200):
        lines.append(" ".join(repr(float(x)) for x in v))
    lines.append("7.0 8.0")
    lines.append('ZONE T="b", I=2, J=2, K=2, DATAPACKING=BLOCK, VARLOCATION=([1-3]=NODAL,[4-4]=CELLCENTERED)')
    for v in range(3):
        lines.append(" ".join(["1.5"] * 8))
    lines.append("9.0")
    p.write_text("\n".join(lines) + "\n")
    zones = caseio.read_tecplot_block(str(p))
    assert len(zones) == 2
    names, I, J, K, data = zones[0]
    assert (I, J, K) == (3, 2, 2) and names == ["x", "y", "z", "q"]
    assert data["x"].shape == (2, 2, 3) and data["x"][1, 0, 2] == 8.0 and data["y"][0, 1, 0] == 103.0
    assert data["q"].shape == (1, 1, 2) and list(data["q"].ravel()) == [7.0, 8.0]
    assert zones[1][4]["q"].ravel().tolist() == [9.0]


def none_face_box():
    """A viscous 12 x 10 x 8 box: periodic in x, adiabatic no-slip walls in y, MOSE 0 (no flux) faces in z, with a
    seeded perturbed state.  Walls and MOSE 0 faces carry no mass flux, so the mass in the box is conserved."""
    from mosegpu import BC_WALL_ADIABATIC, cases
    faces = {1: "periodic", 2: "periodic", 3: BC_WALL_ADIABATIC, 4: BC_WALL_ADIABATIC, 5: BC_NONE, 6: BC_NONE}
    case = cases.box_case((12, 10, 8), faces, extent=(1.0, 0.8, 0.6), limiter="vanleer", riemann="AUSM+", cfl=0.4)
    rng = np.random.Generator(np.random.PCG64(31))
    V = case.layout.interior(case.P, 0)
    sh = V.shape[1:]
    rho, p = 1.1, 1.0e5
    U = 0.3 * np.sqrt(1.4 * p / rho)
    V[0] = rho * (1 + 0.03 * rng.uniform(-1, 1, sh))
    V[1] = U * (1 + 0.2 * rng.uniform(-1, 1, sh))
    V[2] = 0.1 * U * rng.uniform(-1, 1, sh)
    V[3] = 0.1 * U * rng.uniform(-1, 1, sh)
    V[4] = p * (1 + 0.03 * rng.uniform(-1, 1, sh))
    from mosegpu.cases import fill_ghosts_nearest
    fill_ghosts_nearest(case.layout, case.P)
    return case


def test_none_faces_conserve_mass():
    c = none_face_box()
    vol = c.layout.interior(c.geom.vol[None, :], 0)[0]
    m0 = np.sum(c.layout.interior(c.P, 0)[0] * vol)
    P = c.P.copy()
    OracleCase(c, nthreads=8).step(P, 5)
    m1 = np.sum(c.layout.interior(P, 0)[0] * vol)
    assert abs(m1 - m0) <= 1e-13 * m0
    # and the faces are not inert: the state changed
    assert np.max(np.abs(c.layout.interior(P, 0)[1] - c.layout.interior(c.P, 0)[1])) > 0.0


def test_e23_15e3_matches_fortran(tmp_path):
    """The output formatter against Fortran's own E23.15E3 (amdflang), on seeded values over the whole range."""
    import subprocess
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if fc is None:
        pytest.skip("no Fortran compiler")
    rng = np.random.Generator(np.random.PCG64(5))
    vals = np.concatenate([rng.uniform(-1, 1, 300) * 10.0 ** rng.integers(-300, 300, 300), rng.uniform(0, 1, 100),
                           [0.0, 1.0, -1.0, 0.1, 1 / 3, 5e-324, 1.7976931348623157e308, 2.5, 0.05, 9.9999999999999995]])
    (tmp_path / "v.txt").write_text("\n".join(repr(float(v)) for v in vals) + "\n")
    (tmp_path / "f.f90").write_text(
        "program f\n  implicit none\n  real(8) :: v\n  integer :: ios\n  open(10, file='v.txt')\n  do\n"
        "    read(10, *, iostat=ios) v\n    if (ios /= 0) exit\n    write(*, '(E23.15E3)') v\n  end do\nend program\n")
    subprocess.run([fc, "-O0", "f.f90", "-o", "f"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./f"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.splitlines()
    assert out == [caseio.fortran_e23_15e3(v) for v in vals]
    # E20.10, the residual history's descriptor: two exponent digits after E, three in its place beyond 99
    (tmp_path / "f.f90").write_text((tmp_path / "f.f90").read_text().replace("(E23.15E3)", "(E20.10)"))
    subprocess.run([fc, "-O0", "f.f90", "-o", "f"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./f"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout.splitlines()
    assert out == [caseio.fortran_e20_10(v) for v in vals]


def test_residual_history_line():
    """(I8,E20.10,nres E20.10) of Wrap_Setup.f90:140 for nres = 5."""
    ln = caseio.residual_history_line(1234, 0.5, [1.0, 0.0, 2.5e-3, 1e-120, 3.0])
    assert ln == ("    1234" + "    0.5000000000E+00" + "    0.1000000000E+01" + "    0.0000000000E+00"
                  + "    0.2500000000E-02" + "    0.1000000000-119" + "    0.3000000000E+01")


def test_field_tec_layout(tmp_path, vortex):
    c = vortex
    P = c.P.copy()
    p = tmp_path / "field.tec"
    caseio.write_field_tec(c, P, str(p), time=0.25)
    lines = p.read_text().splitlines()
    assert lines[0] == ' VARIABLES ="x" "y" "z" "rho(1)" "u" "v" "w" "p" "T"'
    assert lines[1] == (" ZONE  T = Block1, I=65, J=65, K=2, DATAPACKING=BLOCK, "
                        "VARLOCATION=([1-3]=NODAL,[4-9]=CELLCENTERED), SOLUTIONTIME=0.250000000000000E+000")
    assert len(lines) == 2 + 3 * 65 * 65 * 2 + 6 * 64 * 64
    assert all(len(ln) == 23 for ln in lines[2:])
    names, I, J, K, data = caseio.read_tecplot_block(str(p))[0]
    assert (I, J, K) == (65, 65, 2)
    V = c.layout.interior(P, 0)
    for a, nm in enumerate("xyz"):
        assert np.allclose(data[nm], c.nodes[0][..., a], rtol=1e-14, atol=0)
    for s, nm in enumerate(["rho(1)", "u", "v", "w", "p"]):
        assert np.allclose(data[nm], V[s], rtol=1e-14, atol=0)
    assert np.allclose(data["T"], V[4] / (V[0] * R_AIR), rtol=1e-14, atol=0)
    caseio.write_field_tec(c, P, str(p), time=0.0, time_accurate=False, iteration=7)
    assert p.read_text().splitlines()[1].endswith("SOLUTIONTIME=0.700000000000000E+001, STRANDID = 0")


def test_native_roundtrip(tmp_path, vortex):
    c = vortex
    P = c.P.copy()
    P[1] += 1e-3 * np.sin(np.arange(P.shape[1]))
    p = tmp_path / "field.bin"
    caseio.write_native(c, P, str(p), time=1.5)
    raw = p.read_bytes()
    # record 1: nb = 1 between 4-byte big-endian markers
    assert raw[:12] == (4).to_bytes(4, "big") + (1).to_bytes(4, "big") + (4).to_bytes(4, "big")
    Q, t = caseio.read_native(c, str(p))
    assert t == 1.5
    assert np.array_equal(c.layout.interior(Q, 0), c.layout.interior(P, 0))


SHU_OSHER = os.path.join(HERE, "golden", "cases", "shu-osher")


def test_shu_osher_inputs_and_shock_speed():
    """The reference's test/1D/Shu-Osher inputs: 200 cells on [-5, 5], HLLC + van Leer, RK3 at CFL 0.5, extrapolation
    faces in x.  A Mach 3 shock starts at x = -4 into gas at p = 1, rho = 1 + 0.2 sin 5x; it runs at 3 sqrt(1.4) on
    average (rho 1 in the mean), so after t ~ 1.8 it stands at -4 + 3 sqrt(1.4) t within two cells."""
    c = caseio.load_case(SHU_OSHER, 330)
    s = c.scheme
    assert c.layout.dims == [(200, 1, 1)] and (s.riemann, s.limiter, s.n_rk, s.cfl, s.ndir) == (1, 2, 3, 0.5, 1)
    assert sorted(set(c.records["typ"].tolist())) == [BC_NONE, 11]
    P = c.P.copy()
    t = OracleCase(c).step(P, 330).sum()
    W = c.layout.interior(P, 0)[:, 0, 0, :]
    x = -5.0 + (np.arange(200) + 0.5) * 0.05
    jump = x[np.max(np.where(np.abs(np.diff(W[0])) > 0.5)[0])] + 0.025
    assert 1.7 < t < 1.9
    assert abs(jump - (-4.0 + 3.0 * np.sqrt(1.4) * t)) <= 0.1, (jump, t)
    assert W[0].min() > 0.7 and W[4].min() > 0.9


def test_run_state_sidecar(tmp_path):
    """The run-state sidecar of a native restart (Write_Restart_State / Read_Restart_State): 40 big-endian bytes,
    version 2, and the counters come back only for the field of the same time, to the bit."""
    from mosegpu import caseio
    field = str(tmp_path / "field.bin")
    path = caseio.write_run_state(field, 0.125, 340, 341, 2, 0.0625, 7)
    assert path.endswith("field_state_r0.bin")
    raw = open(path, "rb").read()
    assert len(raw) == 40
    assert raw[:4] == b"\x00\x00\x00\x02" and raw[-4:] == b"\x00\x00\x00\x00"      # version, no multirate partition
    assert np.frombuffer(raw, ">f8", 1, 4)[0] == 0.125
    st = caseio.read_run_state(field, 0.125)
    assert st == {"version": 2, "time": 0.125, "iteration": 340, "iter_general": 341, "n_tdump": 2,
                  "time_from_call": 0.0625, "output_index": 7}
    assert caseio.read_run_state(field, np.nextafter(0.125, 1.0)) is None
    assert caseio.read_run_state(str(tmp_path / "other.bin"), 0.125) is None

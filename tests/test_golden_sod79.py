"""The data of the reference's own Sod test case (test/1D/Sod79/INPUT, committed as tests/golden/sod79_inputs.json
by tests/golden/make_sod79_fixture.py) against the tables and the shock-tube case of this package."""
import json
from pathlib import Path

import numpy as np

from exact_riemann import sample
from mosegpu import cases
from oracle_bind import OracleCase

FIX = json.loads((Path(__file__).resolve().parent / "golden" / "sod79_inputs.json").read_text())


def test_air_table_rows_match_the_reference_thermo_file():
    assert cases.WM_AIR == FIX["wm_air"]
    t = cases.air_tables()
    for T, row in FIX["thermo_rows"].items():
        assert t.cp[0, int(T)] == row["cp"]
        assert t.h[0, int(T)] == row["h"]
    # FLINT's row 0 (Tmin = 0 when the file starts at 1 K): cp(0) = cp(1), h(0) = 0
    assert t.cp[0, 0] == FIX["thermo_rows"]["1"]["cp"] and t.h[0, 0] == 0.0


def test_reference_initial_state_runs_to_the_exact_solution():
    n = FIX["cells"]
    c = cases.sod(n, 0, limiter="mc", riemann="HLLC", cfl=0.9)
    V = c.layout.interior(c.P, 0)
    left = c.meta["x"] < 0.5
    rl, rr = FIX["ic"]["0"]["rho1"], FIX["ic"][str(n - 1)]["rho1"]
    pl, pr = FIX["ic"]["0"]["p"], FIX["ic"][str(n - 1)]["p"]
    assert FIX["ic"][str(n // 2 - 1)]["rho1"] == rl and FIX["ic"][str(n // 2)]["rho1"] == rr
    V[0, 0, 0, :] = np.where(left, rl, rr)
    V[4, 0, 0, :] = np.where(left, pl, pr)
    cases.fill_ghosts_nearest(c.layout, c.P)
    o = OracleCase(c)
    P = c.P.copy()
    t = 0.0
    while t < 0.2 - 1e-15:      # time-threshold = 0.20 of test/1D/Sod79/input.ini
        t += o.step(P, 1)[0]
    g = c.meta["gamma"]
    ex = sample(c.meta["x"], t, 0.5, rl, 0.0, pl, rr, 0.0, pr, g)
    W = c.layout.interior(P, 0)[:, 0, 0, :]
    assert np.mean(np.abs(W[0] - ex[0])) < 2.5 / n
    assert np.mean(np.abs(W[4] - ex[2])) < 2.5 / n

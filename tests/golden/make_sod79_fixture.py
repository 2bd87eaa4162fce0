"""Writes tests/golden/sod79_inputs.json from the data files of the reference's Sod test case
(test/1D/Sod79/INPUT/{phase.txt, thermo.dat, ic.tec}): the molecular weight, sample rows of the thermo table and
the initial cell states at the two ends and around the diaphragm.  Run in a container that has the reference at
/root/reference; the fixture (data only) is what the tests read."""
import json
from pathlib import Path

SRC = Path("/root/reference/test/1D/Sod79/INPUT")
OUT = Path(__file__).resolve().parent / "sod79_inputs.json"


def main():
    phase = SRC.joinpath("phase.txt").read_text().split()
    wm = float(phase[phase.index("air") + 1])
    rows = {}
    for line in SRC.joinpath("thermo.dat").read_text().splitlines():
        w = line.split()
        if len(w) == 5 and w[0].replace(".", "", 1).isdigit():
            T = float(w[0])
            if T in (1.0, 2.0, 300.0, 301.0, 1000.0, 5000.0, 9999.0, 10000.0):
                rows[str(int(T))] = {"cp": float(w[1]), "h": float(w[2])}
    lines = SRC.joinpath("ic.tec").read_text().splitlines()
    hdr = lines[1]
    I = int(hdr.split("I=")[1].split(",")[0])
    J = int(hdr.split("J=")[1].split(",")[0])
    K = int(hdr.split("K=")[1].split(",")[0])
    vals = [float(x) for ln in lines[2:] for x in ln.split()]
    nn, nc = I * J * K, (I - 1) * (J - 1) * (K - 1)
    x = vals[0:nn]
    cell = {}
    base = 3 * nn
    names = ["rho1", "u", "v", "w", "p"]
    for q, nm in enumerate(names):
        cell[nm] = vals[base + q * nc: base + (q + 1) * nc]
    pick = [0, 1, nc // 2 - 2, nc // 2 - 1, nc // 2, nc // 2 + 1, nc - 1]
    out = {"source": "test/1D/Sod79/INPUT (phase.txt, thermo.dat, ic.tec)", "wm_air": wm, "thermo_rows": rows,
           "cells": I - 1, "x_first_nodes": x[:3], "x_last_node": x[I - 1],
           "ic": {str(c): {nm: cell[nm][c] for nm in names} for c in pick}}
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()

"""Case directories in the reference's own input formats (test/<dim>/<case>/{input.ini, INPUT/*}), read into a Case
that this library runs, and its solution written in the reference's output formats (OUTPUT/field.tec, field.bin).

  input.ini      the MOSE_open dialect: [MOSE-Physics], [MOSE-Numerics], [MOSE-Parameters] (IO_Ini_Open.f90:60-320)
  phase.txt      a header line, then one "name molecular-weight" line per species (Load_ThermoTransport.f90:7-60)
  thermo.dat     Tecplot POINT, one zone per species: T, cp, h, s, dcp on T = Tmin..Tmax; with Tmin = 1, row 0 is
                 cp(1) with h = 0 (Load_ThermoTransport.f90:77-101)
  transport.dat  Tecplot POINT, one zone per species: T, mu, k; row 0 = row 1 (Load_ThermoTransport.f90:144-189)
  ic.tec         Tecplot BLOCK, one zone per block: the nodal x, y, z and the cell-centred primitives
  bc.txt         one record per boundary cell-face, 3-digit MOSE_open types (IO_BC_Open.f90:132-399)

Only what this path runs is accepted; anything else is refused by name (local time stepping, the preconditioned
update integration-variables = prec, Riemann solvers other than HLLC / HLLE / SLAU / AUSM+ / LMRoe, boundary types
outside 0/100, 101/201, 300 (symmetry), 301 with zero heat flux, 302, 400 and 405 (the supersonic inlet),
axisymmetric meshes).  The primitive
update (integration-variables = prim) and the implicit residual smoothing (irs, irs-beta) are read into the scheme.  Files may be stored gzip-compressed
(name + ".gz").  write_fortran_case / read_fortran_result are the raw case and result files of the Fortran driver
mose_case_run.
"""
from __future__ import annotations

import gzip
import math
import re
import os

import numpy as np

from .cases import Case, fill_ghosts_nearest, make_scheme
from .core import (BC_CONNECTION, BC_EXTRAPOLATION, BC_FARFIELD, BC_NONE, BC_SYMMETRY, BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL,
                   FARFIELD_DTYPE, LIM, NTAB, RECORD_DTYPE, RS, Layout, ThermoTables, metrics_build)

RUNIV = 8314.46261815324     # Mod_Convective.f90:30


class CaseFormatError(ValueError):
    """A case file that this path cannot run, or that does not parse."""


def _open(path: str):
    if os.path.exists(path):
        return open(path, "rt")
    if os.path.exists(path + ".gz"):
        return gzip.open(path + ".gz", "rt")
    raise FileNotFoundError(path)


# ------------------------------------------------------------------------------------------------- input.ini
def read_ini(path: str) -> dict:
    """{section: {key: value}}; ';' and '#' start comments; keys keep their case, values are stripped."""
    out: dict = {}
    sec = None
    with _open(path) as f:
        for line in f:
            s = line.split(";", 1)[0].split("#", 1)[0].strip()
            if not s:
                continue
            if s.startswith("[") and s.endswith("]"):
                sec = s[1:-1].strip()
                out.setdefault(sec, {})
            elif "=" in s and sec is not None:
                k, v = s.split("=", 1)
                out[sec][k.strip()] = v.strip()
    return out


def _truth(v: str) -> bool:
    return v.strip().lower() in ("true", "t", ".true.", "yes", "1")


# ------------------------------------------------------------------------------------------------- phase, tables
def read_phase(path: str):
    """(names, molecular weights) of phase.txt: the first line is a header."""
    names, wm = [], []
    with _open(path) as f:
        next(f)
        for line in f:
            w = line.split()
            if len(w) >= 2:
                names.append(w[0])
                wm.append(float(w[1]))
    return names, np.array(wm)


def read_tecplot_point(path: str) -> list[np.ndarray]:
    """The zones of a Tecplot POINT file, each as a (rows, nvars) array."""
    zones: list[list[list[float]]] = []
    with _open(path) as f:
        for line in f:
            s = line.strip()
            if not s:
                continue
            head = s.split(None, 1)[0].upper()
            if head.startswith("ZONE"):
                zones.append([])
                continue
            try:
                row = [float(t) for t in s.replace(",", " ").split()]
            except ValueError:
                continue                          # TITLE, VARIABLES, "I=..., F=POINT"
            if zones:
                zones[-1].append(row)
    return [np.array(z, dtype=np.float64) for z in zones]


def _rows_from_zero(T: np.ndarray, col: np.ndarray, what: str, row0_h_zero: bool = False) -> np.ndarray:
    """A table column on T = Tmin..Tmax as this library's rows T = 0..NTAB-1.  Tmin must be 1 (row 0 is then
    filled as FLINT does); rows above Tmax repeat the last row, which a lookup reaches only above Tmax - 1 K."""
    Tmin = int(round(T[0]))
    if Tmin != 1 or np.any(np.diff(np.round(T)) != 1):
        raise CaseFormatError(f"{what}: rows must run T = 1, 2, ... K (got {T[0]} .. {T[-1]})")
    n = T.size
    if n > NTAB - 1:
        raise CaseFormatError(f"{what}: {n} rows; this library's tables end at {NTAB - 1} K")
    out = np.empty(NTAB)
    out[1:n + 1] = col
    out[n + 1:] = col[-1]
    out[0] = 0.0 if row0_h_zero else col[0]
    return out


def read_tables(input_dir: str, viscous: bool) -> ThermoTables:
    names, wm = read_phase(os.path.join(input_dir, "phase.txt"))
    zones = read_tecplot_point(os.path.join(input_dir, "thermo.dat"))
    if len(zones) != len(names):
        raise CaseFormatError(f"thermo.dat has {len(zones)} zones for {len(names)} species")
    cp = np.stack([_rows_from_zero(z[:, 0], z[:, 1], "thermo.dat") for z in zones])
    h = np.stack([_rows_from_zero(z[:, 0], z[:, 2], "thermo.dat", row0_h_zero=True) for z in zones])
    mu = k = None
    if viscous:
        tz = read_tecplot_point(os.path.join(input_dir, "transport.dat"))
        if len(tz) != len(names):
            raise CaseFormatError(f"transport.dat has {len(tz)} zones for {len(names)} species")
        mu = np.stack([_rows_from_zero(z[:, 0], z[:, 1], "transport.dat") for z in tz])
        k = np.stack([_rows_from_zero(z[:, 0], z[:, 2], "transport.dat") for z in tz])
    return ThermoTables(wm, cp, h, mu, k, names=names)


# ------------------------------------------------------------------------------------------------- ic.tec
def _parse_varlocation(header: str, nvar: int) -> list[bool]:
    """True = cell-centred, from VARLOCATION=([a-b]=NODAL,[c-d]=CELLCENTERED); NODAL where absent."""
    cc = [False] * nvar
    up = header.upper()
    p = up.find("VARLOCATION")
    if p < 0:
        return cc
    body = up[up.find("(", p) + 1:up.find(")", p)]
    for rng, loc in re.findall(r"\[([\d,\- ]+)\]\s*=\s*(\w+)", body):
        for piece in rng.split(","):
            a, _, b = piece.partition("-")
            lo, hi = int(a), int(b or a)
            for q in range(lo, hi + 1):
                cc[q - 1] = loc.strip().startswith("CELL")
    return cc


def _zone_int(header: str, key: str) -> int:
    """The integer of `key=` in a ZONE header (a whole key: I= does not match inside another word); 1 if absent."""
    m = re.search(r"(?:^|[,\s])" + key + r"\s*=\s*(\d+)", header, re.IGNORECASE)
    return int(m.group(1)) if m else 1


def read_tecplot_block(path: str):
    """Zones of a Tecplot BLOCK file: [(names, I, J, K, {name: array})]; nodal arrays have shape (K, J, I),
    cell-centred ones (K-1, J-1, I-1) (i fastest, as Tecplot writes them)."""
    with _open(path) as f:
        text = f.read()
    lines = text.splitlines()
    names: list[str] = []
    zones = []
    q = 0
    while q < len(lines):
        s = lines[q].strip()
        up = s.upper()
        if up.startswith("VARIABLES"):
            names = [t for t in s.split("=", 1)[1].replace(",", " ").split()]
            names = [t.strip('"') for t in names]
            q += 1
        elif up.startswith("ZONE"):
            header = s
            q += 1
            # header continuation lines (no leading number)
            while q < len(lines):
                t = lines[q].strip()
                try:
                    float(t.split()[0])
                    break
                except (ValueError, IndexError):
                    header += " " + t
                    q += 1
            I, J, K = _zone_int(header, "I"), _zone_int(header, "J"), _zone_int(header, "K")
            cc = _parse_varlocation(header, len(names))
            counts = [((I - 1) * max(J - 1, 1) * max(K - 1, 1)) if c else I * J * K for c in cc]
            need = sum(counts)
            vals: list[float] = []
            while len(vals) < need and q < len(lines):
                t = lines[q].strip()
                if t:
                    vals.extend(float(x) for x in t.split())
                q += 1
            if len(vals) < need:
                raise CaseFormatError(f"{path}: zone with {len(vals)} values, {need} expected")
            arr = np.array(vals[:need])
            data, o = {}, 0
            for nm, c, n in zip(names, cc, counts):
                shape = (max(K - 1, 1), max(J - 1, 1), I - 1) if c else (K, J, I)
                data[nm] = arr[o:o + n].reshape(shape)
                o += n
            zones.append((names, I, J, K, data))
        else:
            q += 1
    return zones


# ------------------------------------------------------------------------------------------------- bc.txt
# 3-digit MOSE_open types this path runs -> this library's record types (IO_BC_Open.f90:155-211)
_OPEN_TYPES = {0: BC_NONE, 100: BC_NONE, 101: BC_CONNECTION, 201: BC_CONNECTION, 300: BC_SYMMETRY, 301: BC_WALL_ADIABATIC,
               302: BC_WALL_ISOTHERMAL, 400: BC_EXTRAPOLATION, 405: BC_FARFIELD}
# IO_BC_Open.f90:40: the pamb marker of a 405 record, which BC_Inflow_Outflow (Lib_BC.f90:1574-1583) sends to the
# supersonic closure at any Mach and closes on outward flow
PAMB_FORCE_SUPERSONIC = -2.71


def read_bc_open(path: str, inlets: list | None = None) -> np.ndarray:
    """Records of a MOSE_open bc.txt, in file order, as RECORD_DTYPE with 0-based block ids.  A first line
    "b i j k f type" is followed by the type's property line: 101/201 "bs is js ks fs d11 d12 d21 d22",
    301 "qw ks eps", 302 "Tw ks eps", 405 "mach T p alpha beta rel_fac ci(species) ci(rans)" (IO_BC_Open.f90:286-300);
    0, 100, 300 and 400 have none.  The tokens of each 405 line are appended to inlets as (record index, tokens):
    its angles may name the face normal, which needs the metrics (inlet_rows)."""
    recs = []
    with _open(path) as f:
        lines = [ln for ln in (x.strip() for x in f) if ln]
    q = 0
    while q < len(lines):
        w = lines[q].split()
        q += 1
        b, i, j, k, fc, typ = (int(x) for x in w[:6])
        r = np.zeros(1, dtype=RECORD_DTYPE)[0]
        r["b"] = b - 1; r["i"] = i; r["j"] = j; r["k"] = k; r["f"] = fc
        r["d11"] = 1; r["d22"] = 1
        if typ not in _OPEN_TYPES:
            raise CaseFormatError(f"bc.txt: boundary type {typ} (block {b}, face {fc}) is not on this path")
        r["typ"] = _OPEN_TYPES[typ]
        if typ in (101, 201):
            v = [int(x) for x in lines[q].split()[:9]]
            q += 1
            r["bs"] = v[0] - 1; r["is_"] = v[1]; r["js"] = v[2]; r["ks"] = v[3]; r["fs"] = v[4]
            r["d11"], r["d12"], r["d21"], r["d22"] = v[5:9]
            r["periodic"] = 1 if typ == 201 else 0      # 201: ghost nodes translated by one period (metrics)
        elif typ in (301, 302):
            v = [float(x.rstrip(",")) for x in lines[q].split()[:3]]
            q += 1
            if any(x != 0.0 for x in v[1:]):
                raise CaseFormatError("bc.txt: wall roughness / emissivity extras are not on this path")
            if typ == 301 and v[0] != 0.0:
                raise CaseFormatError("bc.txt: a 301 wall with a nonzero heat flux is not on this path")
            if typ == 302:
                r["tw"] = v[0]
        elif typ == 405:
            tok = [x for x in re.split(r"[,\s]+", lines[q]) if x]
            q += 1
            if inlets is not None:
                inlets.append((len(recs), tok))
        recs.append(r)
    return np.array(recs, dtype=RECORD_DTYPE)


# ------------------------------------------------------------------------------------------------- the case
def mesh_ndir(dims, nodes0) -> int:
    """Check_Mesh_Type (Lib_Metrics.f90:139-196) on block 1: 3 if n3 > 1, 1 if n2 = n3 = 1, else 2.  A mesh that
    the reference takes as axisymmetric (2Daxi, or a 3D wedge) is refused."""
    n1, n2, n3 = dims
    nd = nodes0                                      # (n3+1, n2+1, n1+1, 3), node(i, j, k) = nd[k, j, i]
    ang = lambda p: math.atan2(p[2], p[1])
    if n3 > 1:
        t1 = ang(nd[n3, 1, 0]) - ang(nd[0, 1, 0])
        t2 = ang(nd[n3, n2, 0]) - ang(nd[0, n2, 0])
        if (t1 - t2) < 1e-5 and t1 != 0.0:
            raise CaseFormatError("axisymmetric 3D mesh: not on this path")
        return 3
    if n2 == 1:
        return 1
    t1 = ang(nd[1, 1, 0]) - ang(nd[0, 1, 0])
    t2 = ang(nd[1, n2, 0]) - ang(nd[0, n2, 0])
    r0 = math.hypot(nd[0, 1, 0][1], nd[0, 1, 0][2])
    r1 = math.hypot(nd[1, 1, 0][1], nd[1, 1, 0][2])
    if (t1 - t2) < 1e-5 and r0 > 1e-10 and abs(r0 - r1) < 1e-5 * r0 and t1 != 0.0:
        raise CaseFormatError("axisymmetric 2D mesh: not on this path")
    return 2


def scheme_from_ini(ini: dict, ndir: int, nsc: int):
    ph = ini.get("MOSE-Physics", {})
    nu = ini.get("MOSE-Numerics", {})
    eq = ph.get("equations", "")
    if eq not in ("euler", "navier-stokes"):
        raise CaseFormatError(f"equations = {eq!r}: euler or navier-stokes")
    viscous = eq == "navier-stokes"
    turb = ph.get("turbulence", "none") if viscous else "none"
    if turb not in ("none", "SA"):
        raise CaseFormatError(f"turbulence = {turb}: only SA is on this path")
    if not _truth(nu.get("time-accurate", "false")):
        raise CaseFormatError("time-accurate = false (local time stepping) is not on this path")
    rk = {"euler": 1, "RK2": 2, "RK3": 3}.get(nu.get("time-scheme", ""))
    if rk is None:
        raise CaseFormatError(f"time-scheme = {nu.get('time-scheme')!r}: euler, RK2 or RK3")
    intvar = nu.get("integration-variables", "cons")
    if intvar not in ("cons", "prim"):
        raise CaseFormatError(f"integration-variables = {intvar!r}: cons or prim (the preconditioned update, prec, "
                              "is not on this path)")
    recon = nu.get("space-reconstruction", "")
    if recon not in ("MUSCL", "first-order"):
        raise CaseFormatError(f"space-reconstruction = {recon!r}: MUSCL or first-order")
    lim = "first-order" if recon == "first-order" else nu.get("flux-limiter", "").lower()
    if lim not in LIM:
        raise CaseFormatError(f"flux-limiter = {lim!r}: one of {sorted(LIM)}")
    rs = nu.get("riemann-solver", "HLLC")
    if rs not in ("HLLC", "HLLE", "SLAU", "AUSM+", "LMRoe"):
        raise CaseFormatError(f"riemann-solver = {rs}: HLLC, HLLE, SLAU, AUSM+ or LMRoe on this path")
    mco = float(nu.get("riemann-options-Mco", "1.0"))
    if "cfl" not in nu:
        raise CaseFormatError("[MOSE-Numerics] cfl is required")
    kw = {}
    if viscous:
        for key in ("vnn",):
            if key not in nu:
                raise CaseFormatError(f"[MOSE-Numerics] {key} is required by this loader for a viscous case")
        if "schmidt" not in ph or float(ph["schmidt"]) <= 0.0:
            raise CaseFormatError("a viscous case needs [MOSE-Physics] schmidt > 0 (multicomponent diffusion is not "
                                  "on this path)")
        kw = dict(vnn=float(nu["vnn"]), prandtl=float(ph.get("prandtl", "0")), sc=float(ph["schmidt"]))
    pimp = 1
    if turb == "SA":
        pimp = 1 if _truth(ini.get("MOSE-Turbulence", {}).get("point-implicit", "true")) else 0
    # implicit residual smoothing: on with irs = true or irs-beta > 0, beta = irs-beta, two Jacobi sweeps
    # (IO_Ini_Open.f90:138-146, Lib_IRS.f90:10-14)
    try:
        irs_beta = float(nu.get("irs-beta", "0.0"))
    except ValueError:
        irs_beta = 0.0
    irs_on = irs_beta > 0.0 or _truth(nu.get("irs", "false"))
    return make_scheme(nsc=nsc, nrans=1 if turb == "SA" else 0, limiter=lim, riemann=rs, viscous=viscous, n_rk=rk,
                       ndir=ndir, cfl=float(nu["cfl"]), mco=mco, point_implicit=pimp, integration=intvar,
                       irs_beta=irs_beta if irs_on else None, **kw)


def _fortran_real(tok: str) -> float:
    return float(tok.rstrip(",").replace("d", "e").replace("D", "e"))


def inlet_rows(rec: np.ndarray, inlets: list, geom, layout: Layout, nsc: int, nrans: int) -> np.ndarray:
    """The far-field table of the 405 records (one row per distinct set of values) and each record's row index
    (rec["ff"]).  A row: the Mach number, the static T and p (BC_Supersonic_Inflow takes its p0 and T0 as static
    values, IO_BC_Open.f90:286-300), pamb = PAMB_FORCE_SUPERSONIC, rel_fac, the species mass fractions and the RANS
    values.  An angle token "normal" is the face normal's angle, alpha = atan(n2 / (n1 + 1e-20)),
    beta = atan(n3 / (n1 + 1e-20)) (Resolve_Open_BC_Direction, IO_BC_Open.f90:411-451, with the face of
    Face_Index, Lib_BC.f90:46-74); a number is an angle in radians, used as it is."""
    rows, index = [], {}
    for e, tok in inlets:
        if len(tok) < 6 + nsc + nrans:
            raise CaseFormatError(f"bc.txt: a 405 record needs {6 + nsc + nrans} values, has {len(tok)}")
        r = rec[e]
        b, f = int(r["b"]), int(r["f"])
        n1, n2, n3 = layout.dims[b]
        i, j, k = int(r["i"]), int(r["j"]), int(r["k"])
        dr, m = (f + 1) // 2, f % 2
        if dr == 1:
            fm, NF, fi = geom.fm1, layout.NF1, layout.f1off[b] + (i - m) + (n1 + 1) * ((j - 1) + n2 * (k - 1))
        elif dr == 2:
            fm, NF, fi = geom.fm2, layout.NF2, layout.f2off[b] + (i - 1) + n1 * ((j - m) + (n2 + 1) * (k - 1))
        else:
            fm, NF, fi = geom.fm3, layout.NF3, layout.f3off[b] + (i - 1) + n1 * ((j - 1) + n2 * (k - m))
        nx, ny, nz = fm[fi], fm[NF + fi], fm[2 * NF + fi]
        ang = []
        for t, comp in ((tok[3], ny), (tok[4], nz)):
            ang.append(math.atan(comp / (nx + 1e-20)) if "normal" in t else _fortran_real(t))
        row = np.zeros(1, dtype=FARFIELD_DTYPE)[0]
        row["mach"], row["t0"], row["p0"] = (_fortran_real(x) for x in tok[:3])
        row["pamb"] = PAMB_FORCE_SUPERSONIC
        row["alpha"], row["beta"] = ang
        row["rel_fac"] = _fortran_real(tok[5])
        for s in range(nsc):
            row["ci"][s] = _fortran_real(tok[6 + s])
        for s in range(nrans):
            row["rans"][s] = _fortran_real(tok[6 + nsc + s])
        key = row.tobytes()
        if key not in index:
            index[key] = len(rows)
            rows.append(row)
        rec[e]["ff"] = index[key]
    return np.array(rows, dtype=FARFIELD_DTYPE)


def load_case(case_dir: str, nsteps: int | None = None) -> Case:
    """The case of a reference test directory (input.ini and INPUT/), ready for DeviceCase and the oracle.
    nsteps defaults to [MOSE-Parameters] iter-threshold (0 when the case ends on time-threshold)."""
    ini = read_ini(os.path.join(case_dir, "input.ini"))
    inp = os.path.join(case_dir, "INPUT")
    viscous = ini.get("MOSE-Physics", {}).get("equations") == "navier-stokes"
    tables = read_tables(inp, viscous)
    zones = read_tecplot_block(os.path.join(inp, "ic.tec"))
    dims, nodes, fields = [], [], []
    for names, I, J, K, data in zones:
        n1, n2, n3 = I - 1, max(J - 1, 1), max(K - 1, 1)
        dims.append((n1, n2, n3))
        nd = np.stack([data[names[0]], data[names[1]], data[names[2]]], axis=-1)
        nodes.append(np.ascontiguousarray(nd))
        fields.append([data[nm] for nm in names[3:]])
    ndir = mesh_ndir(dims[0], nodes[0])
    sch = scheme_from_ini(ini, ndir, tables.nsc)
    nprim = tables.nsc + 4 + sch.nrans
    for b, fl in enumerate(fields):
        if len(fl) != nprim:
            raise CaseFormatError(f"ic.tec zone {b + 1}: {len(fl)} cell variables, {nprim} expected")
    layout = Layout(dims)
    inlets = []
    rec = read_bc_open(os.path.join(inp, "bc.txt"), inlets)
    order = np.lexsort((rec["f"], rec["b"]))          # grouped by block, then face; file order inside a group
    rec = np.ascontiguousarray(rec[order])
    geom = metrics_build(layout, nodes, rec, ndir=ndir)
    pos = np.empty_like(order)
    pos[order] = np.arange(len(order))
    ff = inlet_rows(rec, [(int(pos[e]), t) for e, t in inlets], geom, layout, tables.nsc, sch.nrans)
    P = np.zeros((nprim, layout.NC))
    for b, fl in enumerate(fields):
        V = layout.interior(P, b)
        for s in range(nprim):
            V[s] = fl[s]
    fill_ghosts_nearest(layout, P)
    par = ini.get("MOSE-Parameters", {})
    if nsteps is None:
        nsteps = int(float(par.get("iter-threshold", "0")))
    return Case(os.path.basename(os.path.normpath(case_dir)), tables, sch, layout, nodes, rec, geom, P, nsteps,
                meta={"ini": ini, "ndir": ndir, "time_threshold": float(par.get("time-threshold", "0") or 0)},
                farfield=ff)


# ------------------------------------------------------------------------------------------------- output files
def fortran_e23_15e3(x: float) -> str:
    """One value as Fortran's E23.15E3 edit descriptor writes it (ORION's FR8P): a 0.d...d mantissa of 15 digits,
    a signed 3-digit exponent, right-justified in 23 columns."""
    x = float(x)
    if x != x:
        return "NaN".rjust(23)
    if x in (math.inf, -math.inf):
        return ("Infinity" if x > 0 else "-Infinity").rjust(23)
    if x == 0.0:
        body, e = "0." + "0" * 15, 0
    else:
        s = "%.14e" % abs(x)                       # d.dddddddddddddde+XX, correctly rounded to 15 digits
        body, e = "0." + s[0] + s[2:16], int(s[17:]) + 1
    sign = "-" if (x < 0.0 or (x == 0.0 and math.copysign(1.0, x) < 0.0)) else ""
    return (sign + body + "E" + ("+" if e >= 0 else "-") + "%03d" % abs(e)).rjust(23)


def fortran_e20_10(x: float) -> str:
    """One value as Fortran's E20.10 edit descriptor writes it: a 0.d...d mantissa of 10 digits and an exponent of
    two digits after "E" (three digits in place of "E" and the sign's column once |exponent| > 99), right-justified
    in 20 columns."""
    x = float(x)
    if x != x:
        return "NaN".rjust(20)
    if x in (math.inf, -math.inf):
        return ("Infinity" if x > 0 else "-Infinity").rjust(20)
    if x == 0.0:
        body, e = "0." + "0" * 10, 0
    else:
        s = "%.9e" % abs(x)                        # d.ddddddddde+XX, correctly rounded to 10 digits
        body, e = "0." + s[0] + s[2:11], int(s[12:]) + 1
    sign = "-" if (x < 0.0 or (x == 0.0 and math.copysign(1.0, x) < 0.0)) else ""
    ex = ("E" + ("+" if e >= 0 else "-") + "%02d" % abs(e)) if abs(e) <= 99 else (("+" if e >= 0 else "-") + "%03d" % abs(e))
    return (sign + body + ex).rjust(20)


def residual_history_line(iteration: int, time: float, norms) -> str:
    """One row of OUTPUT/residual-history.dat (Wrap_Setup.f90:140, Wrap_Postprocess.f90:69): the iteration as I8, the
    time and the nres = nprim - nsc + 1 residual norms as E20.10."""
    return "%8d" % iteration + fortran_e20_10(time) + "".join(fortran_e20_10(v) for v in norms)


def step_with_history(dc, nsteps: int, res_diter: int, path: str, iter0: int = 0, time0: float = 0.0,
                      newrun: bool = True) -> tuple[int, float]:
    """Step a resident case (a DeviceCase) nsteps whole steps and write its residual history the way the reference's
    explicit loop does (Mod_Explicit.f90:324-358, Wrap_Postprocess.f90:66-70): after every step whose iteration
    number is a multiple of res_diter (the [MOSE-IO] diter-res key, default 1), one row with the iteration, the
    simulation time and the L2 norms of that step's change (mose_case_residual_norm).  The device steps up to
    res_diter steps at a time in between.  iter0 and time0 are the iteration and time of the state the case was
    created with (0 for a new run); newrun = False appends to an existing file (a restart).  Returns the iteration
    and time reached."""
    if nsteps < 0 or res_diter < 1:
        raise ValueError("nsteps >= 0 and res_diter >= 1")
    rep = dc.step(0)
    it, t = iter0 + int(rep.steps), time0 + rep.time
    with open(path, "w" if newrun else "a") as f:
        left = nsteps
        while left > 0:
            k = min(res_diter - it % res_diter, left)
            rep = dc.step(k)
            left -= k
            it, t = iter0 + int(rep.steps), time0 + rep.time
            if it % res_diter == 0:
                f.write(residual_history_line(it, t, dc.residual_norm()) + "\n")
    return it, t


def _str_r8p_no_sign(x: float) -> str:
    """ORION's str(no_sign=.true., n=x) for a real(8): the E23.15E3 string without its first column."""
    s = fortran_e23_15e3(x)
    if x > 0.0:
        s = "+" + s[1:]
    return s[1:].strip()


def _output_names(case: Case) -> list[str]:
    """Output_Solution_Setup (IO_Solution.f90:346-378): rho(s), u, v, w, p, the RANS variable, then T."""
    names = [f'"rho({s + 1})"' for s in range(case.scheme.nsc)] + ['"u"', '"v"', '"w"', '"p"']
    if case.scheme.nrans == 1:
        names.append('"mi_tilde"')
    return names + ['"T"']


def _cell_temperature(case: Case, V: np.ndarray) -> np.ndarray:
    """T = p / (rho R) with rho = sum rho_s and R = sum R_s rho_s / rho, summed in species order (write_vtk_tec)."""
    nsc = case.scheme.nsc
    rho = V[0].copy()
    for s in range(1, nsc):
        rho = rho + V[s]
    R = np.zeros_like(rho)
    for s in range(nsc):
        R = R + (RUNIV / case.tables.wm[s]) * V[s] / rho
    return V[nsc + 3] / (rho * R)


def write_field_tec(case: Case, P: np.ndarray, path: str, time: float, time_accurate: bool = True, iteration: int = 0):
    """OUTPUT/field.tec as write_vtk_tec + tec_write_structured_multiblock (IO_Solution.f90:507-584,
    Lib_Tecplot.f90:71-354) write it in ASCII: the variable line, then per block a zone header and, one value per
    line, the nodal x, y, z and the cell-centred primitives and T.  A steady run writes -iteration as its solution
    time and marks the zone STRANDID = 0 (Wrap_Postprocess.f90:43-50)."""
    names = _output_names(case)
    nvar = len(names)
    stime = time if time_accurate else float(iteration)
    lines = [' VARIABLES ="x" "y" "z" ' + " ".join(names)]
    for b, (n1, n2, n3) in enumerate(case.layout.dims):
        head = (f" ZONE  T = Block{b + 1}, I={n1 + 1}, J={n2 + 1}, K={n3 + 1}, DATAPACKING=BLOCK"
                f", VARLOCATION=([1-3]=NODAL,[4-{3 + nvar}]=CELLCENTERED), SOLUTIONTIME={_str_r8p_no_sign(stime)}")
        if not time_accurate:
            head += ", STRANDID = 0"
        lines.append(head)
        nd = case.nodes[b]
        for a in range(3):
            lines.extend(fortran_e23_15e3(v) for v in nd[..., a].reshape(-1))
        V = case.layout.interior(P, b)
        for s in range(case.nprim):
            lines.extend(fortran_e23_15e3(v) for v in V[s].reshape(-1))
        lines.extend(fortran_e23_15e3(v) for v in _cell_temperature(case, V).reshape(-1))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _record(f, arr: np.ndarray):
    data = np.ascontiguousarray(arr).astype(arr.dtype.newbyteorder(">"), copy=False).tobytes()
    mark = np.array([len(data)], dtype=">i4").tobytes()
    f.write(mark + data + mark)


def write_native(case: Case, P: np.ndarray, path: str, time: float):
    """OUTPUT/field.bin as write_native (IO_Solution.f90:450-500): a big-endian Fortran unformatted sequential file
    (4-byte record markers).  Records: nb; dim(1), dim(2), dim(3) one block per record; the species of each block
    with s fastest; each of u, v, w, p over the blocks; the RANS variables of each block with s fastest; the time."""
    dims = case.layout.dims
    nsc, nrans = case.scheme.nsc, case.scheme.nrans
    with open(path, "wb") as f:
        _record(f, np.array([len(dims)], dtype=np.int32))
        for a in range(3):
            for d in dims:
                _record(f, np.array([d[a]], dtype=np.int32))
        for b in range(len(dims)):
            V = case.layout.interior(P, b)                       # (nprim, n3, n2, n1)
            _record(f, np.moveaxis(V[:nsc], 0, -1).reshape(-1))  # s fastest, then i, j, k
        for s in range(nsc, nsc + 4):
            for b in range(len(dims)):
                _record(f, case.layout.interior(P, b)[s].reshape(-1))
        if nrans > 0:
            for b in range(len(dims)):
                V = case.layout.interior(P, b)
                _record(f, np.moveaxis(V[nsc + 4:], 0, -1).reshape(-1))
        _record(f, np.array([time], dtype=np.float64))


def read_native(case: Case, path: str) -> tuple[np.ndarray, float]:
    """The state and time of a field.bin written by write_native (or by the reference's write_native) on the case's
    blocks; the ghost rows are filled from the nearest interior cell."""
    with open(path, "rb") as f:
        raw = f.read()
    recs, q = [], 0
    while q < len(raw):
        n = int(np.frombuffer(raw, ">i4", 1, q)[0])
        recs.append(raw[q + 4:q + 4 + n])
        q += 8 + n
    dims = case.layout.dims
    nb = int(np.frombuffer(recs[0], ">i4")[0])
    got = [tuple(int(np.frombuffer(recs[1 + a * nb + b], ">i4")[0]) for a in range(3)) for b in range(nb)]
    if nb != len(dims) or [tuple(d) for d in dims] != got:
        raise CaseFormatError(f"{path}: blocks {got} do not match the case's {dims}")
    nsc, nrans = case.scheme.nsc, case.scheme.nrans
    P = np.zeros((case.nprim, case.layout.NC))
    r = 1 + 3 * nb
    for b, (n1, n2, n3) in enumerate(dims):
        sp = np.frombuffer(recs[r], ">f8").reshape(n3, n2, n1, nsc)
        case.layout.interior(P, b)[:nsc] = np.moveaxis(sp, -1, 0)
        r += 1
    for s in range(nsc, nsc + 4):
        for b, (n1, n2, n3) in enumerate(dims):
            case.layout.interior(P, b)[s] = np.frombuffer(recs[r], ">f8").reshape(n3, n2, n1)
            r += 1
    if nrans > 0:
        for b, (n1, n2, n3) in enumerate(dims):
            tv = np.frombuffer(recs[r], ">f8").reshape(n3, n2, n1, nrans)
            case.layout.interior(P, b)[nsc + 4:] = np.moveaxis(tv, -1, 0)
            r += 1
    time = float(np.frombuffer(recs[r], ">f8")[0])
    fill_ghosts_nearest(case.layout, P)
    return P, time


def run_state_path(field_path: str, rank: int = 0) -> str:
    """OUTPUT/<field>_state_r<rank>.bin beside the native restart file <field>.bin (Write_Restart_State,
    Wrap_Postprocess.f90:337-357)."""
    base = field_path[:-4] if field_path.endswith(".bin") else field_path
    return f"{base}_state_r{rank}.bin"


def write_run_state(field_path: str, time: float, iteration: int, iter_general: int | None = None, n_tdump: int = 0,
                    time_from_call: float = 0.0, output_index: int = 0, rank: int = 0) -> str:
    """The run-state sidecar of a native restart file (Write_Restart_State, Wrap_Postprocess.f90:337-357): a
    big-endian stream of the version (2), the time, the iteration, the general iteration, the count of time-based
    dumps, the time of the call and the output file index, then the multirate partition flag (a Fortran logical,
    .false.: no multirate partition on this path, MR_Write_State, Mod_Multirate_Partition.f90:425-446)."""
    path = run_state_path(field_path, rank)
    with open(path, "wb") as f:
        f.write(np.array([2], dtype=">i4").tobytes())
        f.write(np.array([time], dtype=">f8").tobytes())
        f.write(np.array([iteration, iteration if iter_general is None else iter_general, n_tdump], dtype=">i4").tobytes())
        f.write(np.array([time_from_call], dtype=">f8").tobytes())
        f.write(np.array([output_index, 0], dtype=">i4").tobytes())
    return path


def read_run_state(field_path: str, time: float, rank: int = 0) -> dict | None:
    """The counters of a run-state sidecar, applied only when its time is the restart field's time to the bit
    (Read_Restart_State, Wrap_Setup.f90:318-364): None when the file is absent or of another time."""
    path = run_state_path(field_path, rank)
    if not os.path.exists(path):
        return None
    raw = open(path, "rb").read()
    ver = int(np.frombuffer(raw, ">i4", 1, 0)[0])
    t = float(np.frombuffer(raw, ">f8", 1, 4)[0])
    it, itg, ntd = (int(x) for x in np.frombuffer(raw, ">i4", 3, 12))
    tfc = float(np.frombuffer(raw, ">f8", 1, 24)[0])
    ids = int(np.frombuffer(raw, ">i4", 1, 32)[0]) if ver >= 2 else -1
    if t != time:
        return None
    return {"version": ver, "time": t, "iteration": it, "iter_general": itg, "n_tdump": ntd, "time_from_call": tfc,
            "output_index": ids}


def run_case(dc, case):
    """Runs a loaded reference case to its end on a DeviceCase: iter-threshold steps when the case gives them, else
    until its time-threshold (DeviceCase.run_until).  Returns the last step report."""
    if case.nsteps > 0:
        return dc.step(case.nsteps)
    t_end = case.meta.get("time_threshold", 0.0)
    if not t_end > 0.0:
        raise CaseFormatError("the case gives neither iter-threshold nor time-threshold")
    return dc.run_until(t_end)


def write_fortran_case(case, path, nsteps: int, tune_path: int = 0, tile=(0, 0, 0), mech=None, opts=None):
    """The raw case file of the Fortran driver mose_case_run (fortran/mose_case_run.f90 gives the layout): the scheme,
    blocks, geometry, records, far-field rows, tables and initial state of `case`, the step count and tuning, and for a
    reacting case the mechanism (mosegpu.chem.Mechanism) with the integrator options (chem.ode_opts)."""
    t = case.tables
    visc_tabs = int(t.mu is not None and t.k is not None)
    nz = case.scheme.nsc + 1
    with open(path, "wb") as f:
        hdr = [1, case.layout.nb, len(case.records), len(case.farfield), nsteps, tune_path, tile[0], tile[1], tile[2],
               int(mech is not None), 0 if opts is None else opts.method, visc_tabs]
        np.array(hdr, dtype=np.int32).tofile(f)
        f.write(bytes(case.scheme))
        np.ascontiguousarray(case.layout.dims_arr, dtype=np.int32).tofile(f)
        for a in (t.wm, t.cp, t.h) + ((t.mu, t.k) if visc_tabs else ()):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        lay, g = case.layout, case.geom
        np.array([lay.NC, lay.NF1, lay.NF2, lay.NF3], dtype=np.int64).tofile(f)
        for a in (g.dl, g.m, g.vol, g.yn, g.fm1, g.fm2, g.fm3):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        np.ascontiguousarray(case.records).tofile(f)
        np.ascontiguousarray(case.farfield).tofile(f)
        np.ascontiguousarray(case.P, dtype=np.float64).tofile(f)
        if mech is not None:
            np.array([mech.ns, mech.n_arrh, mech.n_troe, mech.n_lind, 0, mech.cp.shape[1] - 1], dtype=np.int32).tofile(f)
            for a in (mech.wm, mech.cp, mech.h, mech.ni1, mech.ni2, mech.eps, mech.kf, mech.kb, mech.troe, mech.lind):
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)
            np.array([opts.rtol[q] for q in range(nz)], dtype=np.float64).tofile(f)
            np.array([opts.atol[q] for q in range(nz)], dtype=np.float64).tofile(f)


def read_fortran_result(case, path):
    """(P, time, steps) of mose_case_run's output file."""
    raw = np.fromfile(path, dtype=np.uint8)
    n = case.nprim * case.layout.NC
    P = raw[:8 * n].view(np.float64).reshape(case.nprim, case.layout.NC).copy()
    time = float(raw[8 * n:8 * n + 8].view(np.float64)[0])
    steps = int(raw[8 * n + 8:8 * n + 16].view(np.int64)[0])
    return P, time, steps

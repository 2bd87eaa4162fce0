"""Helpers of the Fortran case-directory tests: a reference case directory with its gzip fixtures expanded (the
Fortran host reads plain files), the mose_case_dir program, and its dump read back as arrays."""
import gzip
import os
import shutil
import subprocess

import numpy as np

import mosegpu
from mosegpu.core import FARFIELD_DTYPE, NTAB, RECORD_DTYPE, Scheme

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "cases")
CASES = ("isentropic-vortex", "shu-osher", "ein91", "noh87", "diff1", "sod79", "forward-step")
PROGRAM = mosegpu.LIB_PATH.parent / "mose_case_dir"


def expand(name: str, dest) -> str:
    """The fixture directory of `name` with every .gz file expanded, under dest; returns the case directory."""
    src = os.path.join(GOLD, name)
    out = os.path.join(str(dest), name)
    for root, _, files in os.walk(src):
        rel = os.path.relpath(root, src)
        os.makedirs(os.path.join(out, rel), exist_ok=True)
        for f in files:
            s = os.path.join(root, f)
            if f.endswith(".gz"):
                with gzip.open(s, "rb") as a, open(os.path.join(out, rel, f[:-3]), "wb") as b:
                    shutil.copyfileobj(a, b)
            else:
                shutil.copy(s, os.path.join(out, rel, f))
    return out


def run(case_dir: str, nsteps: int, out: str, *extra, timeout: int = 300):
    assert PROGRAM.exists(), "build the Fortran host first (make -C mose-gpu_amd)"
    return subprocess.run([str(PROGRAM), case_dir, str(nsteps), out, *[str(x) for x in extra]], capture_output=True,
                          text=True, timeout=timeout)


def read_dump(path: str, viscous: bool, nprim: int) -> dict:
    raw = open(path, "rb").read()
    o = 0

    def take(dt, n):
        nonlocal o
        a = np.frombuffer(raw, dtype=dt, count=int(n), offset=o)
        o += a.nbytes
        return a

    nb, nsc, nrans, nrec, ndir = (int(x) for x in take(np.int32, 5))
    d = {"nb": nb, "nsc": nsc, "nrans": nrans, "ndir": ndir}
    d["scheme"] = bytes(take(np.uint8, len(bytes(Scheme()))))
    d["dims"] = take(np.int32, 3 * nb)
    d["wm"] = take(np.float64, nsc)
    d["cp"] = take(np.float64, NTAB * nsc)
    d["h"] = take(np.float64, NTAB * nsc)
    if viscous:
        d["mu"] = take(np.float64, NTAB * nsc)
        d["k"] = take(np.float64, NTAB * nsc)
    NC, NF1, NF2, NF3 = (int(x) for x in take(np.int64, 4))
    d["dl"] = take(np.float64, 3 * NC)
    d["m"] = take(np.float64, 9 * NC)
    d["vol"] = take(np.float64, NC)
    d["fm1"] = take(np.float64, 4 * NF1)
    d["fm2"] = take(np.float64, 4 * NF2)
    d["fm3"] = take(np.float64, 4 * NF3)
    d["rec"] = take(RECORD_DTYPE, nrec)
    nff = int(take(np.int32, 1)[0])
    d["ff"] = take(FARFIELD_DTYPE, nff)
    d["P"] = take(np.float64, NC * nprim)
    assert o == len(raw), "dump size"
    return d


def read_result(path: str, NC: int, nprim: int):
    raw = open(path, "rb").read()
    P = np.frombuffer(raw, dtype=np.float64, count=NC * nprim).reshape(nprim, NC).copy()
    t = float(np.frombuffer(raw, dtype=np.float64, count=1, offset=8 * NC * nprim)[0])
    steps = int(np.frombuffer(raw, dtype=np.int64, count=1, offset=8 * NC * nprim + 8)[0])
    return P, t, steps

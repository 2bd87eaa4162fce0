#!/usr/bin/env python3
"""Benchmark of the residual path on one MI355X: the 3D compressible Taylor-Green vortex, 256^3 periodic cells,
M = 0.1, Re = 1600, Roe (LMRoe, Mco = 1) + van Albada MUSCL, central viscous fluxes with Sutherland viscosity,
SSP-RK3, FP64.  One "step" = one SSP-RK3 step of all 16.8 M cells (three fused stage kernels).

Prints one JSON line: value = cell-updates/s (cells x steps / s) with the state resident in HBM; roofline of the
fused stage kernel from HIP events on its own stream (algorithmic bytes per cell-stage 8 (3 nprim + 25) = 320 B);
cpu_baseline = the CPU oracle (a C restatement of the reference's double-precision path) on a bounded sample.
Beside the headline line's own fields: "serving" (continuous batching of a Poisson stream of cases, p50/p99
time-to-solution, against the plain per-case path) and "chemistry" (batched RODAS cells/s with its CPU sample, and
RADAU5 beside it).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "mose-gpu_amd"))

METRIC = "FP64 cell-updates/s & %HBM roofline; served cases/s + p50/p99 at ≤1e-12 err"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_cell_stage(nprim: int, viscous: bool, rans: bool) -> int:
    # read the stage input, read the step base, write the stage output (3 nprim) + geometry read once per cell:
    # 12 face normals/areas + 3 lengths + 1 volume (+ 9 metric if viscous, + 1 wall distance if RANS)
    G = 16 + (9 if viscous else 0) + (1 if rans else 0)
    return 8 * (3 * nprim + G)


def profiled_counts(n: int, nprim: int, viscous: bool, tile, path: str = "fused") -> tuple[int | None, int | None]:
    """(HBM bytes, VALU wavefront instructions) per stage launch, measured by rocprofv3 counters for this workload
    (profiles/*_traffic.txt), or None where the workload has not been profiled."""
    kern = "k_stage" if path == "fused" else f"k_stage_{path}"
    key = (kern, str(n), str(nprim), str(int(viscous)), "x".join(str(t) for t in tile))
    for f in sorted((REPO / "profiles").glob("*_traffic.txt")):
        for line in f.read_text().splitlines():
            w = line.split()
            if len(w) >= 6 and not line.startswith("#") and tuple(w[:5]) == key:
                return int(w[5]), (int(w[6]) if len(w) > 6 else None)
    return None, None


# FP64 VALU issue: a wave64 instruction takes 4 cycles of one SIMD; MI355X has 256 CUs x 4 SIMDs at 2.4 GHz
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4


def host_cpu() -> dict:
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def oracle_rate(n: int, steps: int, threads: int) -> tuple[float, float]:
    """cell-updates/s of the CPU oracle on the Taylor-Green case at n^3 over `steps` steps after a warm-up step."""
    sys.path.insert(0, str(REPO / "tests"))
    from mosegpu import cases
    from oracle_bind import OracleCase
    case = cases.taylor_green(n, steps)
    o = OracleCase(case, nthreads=threads)
    P = case.P.copy()
    o.step(P, 1)                       # warm-up step
    t0 = time.perf_counter()
    o.step(P, steps)
    dt = time.perf_counter() - t0
    return n ** 3 * steps / dt, dt


def cpu_baseline(n: int, steps: int, threads: int):
    rate, dt = oracle_rate(n, steps, threads)
    n1, s1 = 40, 32
    rate1, dt1 = oracle_rate(n1, s1, 1)
    return {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"taylor-green {n}^3, {steps} SSP-RK3 steps after 1 warm-up step, CPU oracle "
                      f"(gcc -O3 -march=x86-64-v3 -ffp-contract=off, OpenMP {threads} threads), {dt:.1f} s",
            "single_thread": {"value": rate1, "sample": f"taylor-green {n1}^3, {s1} steps, 1 thread, {dt1:.1f} s"},
            "host": host_cpu()}


def chemistry_leg(ncell: int, cpu_cells: int, threads: int, radau_cells: int = 0, radau_cpu: int = 0):
    """Batched stiff chemistry: independent cells of the 9-species hydrogen-air mechanism at 1 atm, T uniform in
    [900, 2500] K, equivalence ratio uniform in [0.5, 2] (seeded), each integrated over dt = 1e-6 s with RODAS at
    rtol 1e-8 / atol 1e-12 in one persistent launch.  The 2M-cell batch is sampled by `ncell` cells (the per-cell
    work does not depend on the batch size once the device is full); cells/s from the kernel's HIP events.
    `radau_cells` > 0 adds the same cells under RADAU5 (its own launch) as chemistry["radau5"]."""
    import numpy as np
    from mosegpu import Context, DeviceChem, chem
    from mosegpu.core import ThermoTables
    M = chem.h2_air_mechanism()
    st = chem.h2_air_cells(ncell, seed=22)
    ctx = Context(ThermoTables(M.wm, M.cp, M.h, names=M.names))
    dc = DeviceChem(ctx, M, ncell)
    dc.upload(st)
    t0 = time.perf_counter()
    rep = dc.advance(1e-6, chem.ode_opts(1e-8, 1e-12))
    wall = time.perf_counter() - t0
    out, hr, idid, nstep = dc.download()
    dc.destroy()
    ctx.close()
    res = {"cells": ncell, "value": ncell / (rep.ms * 1e-3), "unit": "cells/s", "kernel": "k_chem (RODAS, a team of ns+1 "
           "lanes per cell, persistent queue)", "kernel_ms": rep.ms, "wall_s": wall, "steps_mean": float(nstep.mean()),
           "steps_max": int(nstep.max()), "rhs_evals": int(rep.rhs_total), "failed": int((idid < 0).sum()),
           "workload": "H2-O2-N2 9 species, 17 Arrhenius + 2 Troe, 1 atm, T U[900,2500] K, phi U[0.5,2], "
                       "dt 1e-6 s, RODAS rtol 1e-8 atol 1e-12"}
    per = chem_valu_per_attempt()
    if per:
        # FP64 VALU issue: a wave64 instruction takes 4 cycles of one SIMD (1,024 SIMDs at 2.4 GHz); the share of the
        # launch that the profiled instructions per step attempt, times this launch's attempts, take to issue
        issue_s = per * rep.steps_total / VALU_ISSUE_PER_S      # steps_total counts every attempt (NSTEP)
        res["roofline"] = {"bound": "fp64-valu", "valu_insts_per_step_attempt": per,
                           "step_attempts": int(rep.steps_total),
                           "valu_issue_s": issue_s, "valu_issue_frac": issue_s / (rep.ms * 1e-3)}
    print(f"bench: chemistry {ncell} cells in {rep.ms / 1e3:.2f} s: {res['value']:.3e} cells/s", file=sys.stderr,
          flush=True)
    if radau_cells > 0:
        res["radau5"] = radau5_leg(M, chem.h2_air_cells(radau_cells, seed=22), radau_cpu, threads)
    if cpu_cells > 0:
        sys.path.insert(0, str(REPO / "tests"))
        from oracle_bind import OracleChem
        sample = st[:, :cpu_cells].copy()
        o = OracleChem(M, nthreads=threads)
        t0 = time.perf_counter()
        o.advance(sample, 1e-6)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": cpu_cells / dt, "unit": "cells/s", "cores": threads, "kind": "port",
                               "sample": f"the first {cpu_cells} cells of the same batch, CPU oracle RODAS (gcc -O3 "
                                         f"-march=x86-64-v3, OpenMP {threads} threads), {dt:.1f} s"}
    return res


def radau5_leg(M, st, cpu_cells: int, threads: int):
    """The chemistry cells under RADAU5 (rtol 1e-8 / atol 1e-12).  Cells on which the simplified Newton iteration
    stalls run to NMAX = 100000 steps and end with idid -2, as in the reference's own RADAU5; one such cell takes a
    team the whole launch, so the launch time is set by that tail."""
    from mosegpu import Context, DeviceChem, MoseError, chem
    from mosegpu.core import ThermoTables
    n = st.shape[1]
    ctx = Context(ThermoTables(M.wm, M.cp, M.h, names=M.names))
    dc = DeviceChem(ctx, M, n)
    dc.upload(st)
    try:
        rep = dc.advance(1e-6, chem.ode_opts(1e-8, 1e-12, method=2))
    except MoseError:
        rep = dc.last_report     # cells that stop at NMAX are reported as an error; the launch itself completed
    out, hr, idid, nstep = dc.download()
    dc.destroy()
    ctx.close()
    stalled = int((idid == -2).sum())
    res = {"cells": n, "value": n / (rep.ms * 1e-3), "unit": "cells/s", "kernel": "k_radau (RADAU5, a team of ns+1 "
           "lanes per cell, persistent queue)", "kernel_ms": rep.ms, "steps_mean": float(nstep.mean()),
           "steps_max": int(nstep.max()), "stopped_at_nmax": stalled, "failed_other": int((idid < 0).sum()) - stalled,
           "steps_mean_completed": float(nstep[idid >= 0].mean()) if (idid >= 0).any() else None}
    print(f"bench: radau5 {n} cells in {rep.ms / 1e3:.2f} s: {res['value']:.3e} cells/s, {stalled} stopped at NMAX",
          file=sys.stderr, flush=True)
    if cpu_cells > 0:
        sys.path.insert(0, str(REPO / "tests"))
        from oracle_bind import OracleChem
        sample = st[:, :cpu_cells].copy()
        o = OracleChem(M, nthreads=threads)
        t0 = time.perf_counter()
        o.advance(sample, 1e-6, method=2)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": cpu_cells / dt, "unit": "cells/s", "cores": threads, "kind": "port",
                               "sample": f"the first {cpu_cells} cells of the same batch, CPU oracle RADAU5 (gcc -O3 "
                                         f"-march=x86-64-v3, OpenMP {threads} threads), {dt:.1f} s"}
    return res


def shock_tube_leg(threads: int):
    """The CPU-runnable configuration: the 1D shock tube, 1000 cells, minmod + HLLC, SSP-RK3 at CFL 0.5, 200 steps,
    on the device (fused parity path, the whole run with its launches) and on the CPU oracle (1 thread and all);
    the device's final state must be the oracle's bit for bit.  Launch-bound on the device: 2e5 cell-updates."""
    import numpy as np
    import mosegpu
    from mosegpu import cases
    case = cases.sod(1000, 200)
    ctx = mosegpu.Context(case.tables)
    dc = mosegpu.DeviceCase(ctx, case)
    dc.upload(case.P)
    dc.tune(mosegpu.PATH_FUSED, (0, 0, 0))
    dc.step(0)
    t0 = time.perf_counter()
    dc.step(200)
    el = time.perf_counter() - t0
    Pg = dc.download()
    dc.destroy()
    ctx.close()
    res = {"workload": "shock tube 1000 cells, (1,0,1)|(0.125,0,0.1), minmod + HLLC, SSP-RK3 CFL 0.5, 200 steps",
           "cells": 1000, "steps": 200, "gpu_s": el, "gpu_cell_updates_per_s": 1000 * 200 / el}
    if threads > 0:
        sys.path.insert(0, str(REPO / "tests"))
        from oracle_bind import OracleCase
        for nt in sorted({1, threads}):
            P = case.P.copy()
            o = OracleCase(case, nthreads=nt)
            t0 = time.perf_counter()
            o.step(P, 200)
            dt = time.perf_counter() - t0
            res[f"cpu_{nt}_threads_s"] = dt
            res["same_state_as_oracle"] = bool(np.array_equal(case.layout.interior(P, 0),
                                                              case.layout.interior(Pg, 0)))
    print(f"bench: shock tube 1000 x 200 steps: device {el * 1e3:.1f} ms" +
          "".join(f", CPU {k.split('_')[1]} thread(s) {v * 1e3:.1f} ms" for k, v in res.items()
                  if k.startswith("cpu_")), file=sys.stderr, flush=True)
    return res


def chem_valu_per_attempt() -> float | None:
    """VALU wavefront instructions per RODAS step attempt of k_chem, from the profiled counters
    (profiles/*_traffic.txt key k_chem_valu_per_attempt), or None."""
    for f in sorted((REPO / "profiles").glob("*_traffic.txt")):
        for line in f.read_text().splitlines():
            w = line.split()
            if len(w) >= 2 and w[0] == "k_chem_valu_per_attempt":
                return float(w[1])
    return None


def fast_vs_parity(dc, case, tile, steps: int):
    """The performance path against the parity path (the oracle's bits) on the benchmark's own case: both from the
    initial state for `steps` steps, the relative L-infinity of every primitive variable over the whole box."""
    import numpy as np
    import mosegpu
    out = {}
    for name, path in (("fast", mosegpu.PATH_FAST), ("parity", mosegpu.PATH_FUSED)):
        dc.upload(case.P)
        dc.tune(path, tile, timing=False)
        dc.step(steps)
        out[name] = case.layout.interior(dc.download(), 0).reshape(case.nprim, -1)
    a, b = out["fast"], out["parity"]
    ns = case.scheme.nsc
    scale = np.abs(b).max(axis=1)
    scale[ns:ns + 3] = np.abs(b[ns:ns + 3]).max()        # the velocity as a vector: one scale for its components
    err = float((np.abs(a - b).max(axis=1) / scale).max())
    print(f"bench: fast path vs parity path after {steps} steps: rel Linf {err:.3e}", file=sys.stderr, flush=True)
    return {"steps": steps, "rel_linf": err, "bar": 1e-12, "within_bar": err <= 1e-12}


def lattice_leg(steps: int, warmup: int = 2):
    """The 4 x 4 x 4 lattice of blocks with seeded uneven widths in [24, 96] per axis (cases.channel_lattice, seed 7),
    one-to-one interfaces, an adiabatic no-slip wall at y = 0, far field on the other outer faces, M 0.3, Re 1e6,
    Spalart-Allmaras, LMRoe (Mco = 1) + van Albada, SSP-RK3: the stage kernel over the tiles of every block (the
    performance path, and the parity path) against the plain per-block launch path, on the same start state.  The
    parity path's final state hash must equal the plain path's; the performance path's final state is compared with
    the parity path's (relative L-infinity per variable)."""
    import numpy as np
    import mosegpu
    from mosegpu import PATH_FAST, PATH_FUSED, PATH_PLAIN, cases
    case = cases.channel_lattice(riemann="LMRoe", nsteps=steps)
    cells = case.interior_cells()
    out = {"blocks": len(case.layout.dims), "cells": cells, "steps": steps,
           "workload": "4x4x4 block lattice, widths U{24..96} (seed 7), wall + far field, SA, M 0.3, Re 1e6, "
                       "LMRoe + van Albada, SSP-RK3"}
    hashes, states = {}, {}
    for name, path in (("fast", PATH_FAST), ("fused", PATH_FUSED), ("plain", PATH_PLAIN)):
        ctx = mosegpu.Context(case.tables)
        dc = mosegpu.DeviceCase(ctx, case)
        dc.upload(case.P)
        chosen = dc.tune(path, (0, 0, 0), timing=False)
        dc.step(warmup)
        dc.step(0)
        t0 = time.perf_counter()
        dc.step(steps)
        el = time.perf_counter() - t0
        hashes[name] = dc.state_hash()
        if name != "plain":
            states[name] = np.hstack([case.layout.interior(dc.download(), b).reshape(case.nprim, -1)
                                      for b in range(case.layout.nb)])
        out[name] = {"cell_updates_per_s": cells * steps / el, "ms_per_step": 1e3 * el / steps,
                     "tile": list(chosen) if path != PATH_PLAIN else None}
        dc.destroy()
        ctx.close()
    out["fused_over_plain"] = out["fused"]["cell_updates_per_s"] / out["plain"]["cell_updates_per_s"]
    out["fast_over_plain"] = out["fast"]["cell_updates_per_s"] / out["plain"]["cell_updates_per_s"]
    out["same_state"] = hashes["fused"] == hashes["plain"]
    a, b = states["fast"], states["fused"]
    ns = case.scheme.nsc
    scale = np.abs(b).max(axis=1)
    scale[ns:ns + 3] = float((np.sqrt((b[ns:ns + 3] ** 2).sum(axis=0)) + np.sqrt(b[ns + 3] / b[:ns].sum(axis=0))).max())
    scale[scale == 0] = 1.0
    out["fast_vs_parity_rel_linf"] = float((np.abs(a - b).max(axis=1) / scale).max())
    print(f"bench: lattice {cells} cells in 64 blocks: fast {out['fast']['ms_per_step']:.2f}, fused "
          f"{out['fused']['ms_per_step']:.2f}, plain {out['plain']['ms_per_step']:.2f} ms/step; fused == plain "
          f"{out['same_state']}, fast vs parity {out['fast_vs_parity_rel_linf']:.3e}", file=sys.stderr, flush=True)
    return out


def serving_leg(count: int, plain: bool):
    """A Poisson stream (20 cases/s) of 2D four-quadrant Riemann (n log-uniform in [64, 1024]) and 3D TGV (n
    log-uniform in [48, 192]) cases of 50-500 steps, served with continuous batching on the performance path, and
    again on the parity path; p50/p99 time-to-solution.  The comparison path runs the same requests one by one on the
    plain per-block launch path."""
    from mosegpu import PATH_FAST, PATH_FUSED, PATH_PLAIN, serve
    reqs = serve.request_stream(count, rate=20.0, seed=2024)
    st = serve.serve(reqs, quantum=10, path=PATH_FAST)
    res = {"requests": count, "rate_per_s": 20.0, "path": "fast", "cases_per_s": st.cases_per_s, "p50_s": st.p50,
           "p99_s": st.p99, "wall_s": st.wall, "max_resident": st.max_resident,
           "cell_updates_per_s": st.cell_updates / st.wall,
           "workload": "poisson 20/s: 2D 4-quadrant Riemann n^2 (n in [64,1024]) | 3D TGV n^3 (n in [48,192]), "
                       "50-500 steps"}
    print(f"bench: served {count} cases (performance path) in {st.wall:.2f} s, p50 {st.p50:.3f} s, p99 {st.p99:.3f} s",
          file=sys.stderr, flush=True)
    reqs_p = serve.request_stream(count, rate=20.0, seed=2024)
    sp = serve.serve(reqs_p, quantum=10, path=PATH_FUSED)
    res["parity_path"] = {"cases_per_s": sp.cases_per_s, "p50_s": sp.p50, "p99_s": sp.p99, "wall_s": sp.wall}
    print(f"bench: served {count} cases (parity path) in {sp.wall:.2f} s, p50 {sp.p50:.3f} s, p99 {sp.p99:.3f} s",
          file=sys.stderr, flush=True)
    if plain:
        # the comparison path: each request alone, one after another, on the plain per-block launch path; its
        # whole-request time (host preparation included, as in the served time-to-solution) and its stepping time
        t_req = t_dev = 0.0
        for r in serve.request_stream(count, rate=20.0, seed=2024):
            _, dt, tr = serve.run_alone(r, PATH_PLAIN)
            t_dev += dt
            t_req += tr
        res["plain_one_by_one_cases_per_s"] = count / t_req
        res["speedup_vs_plain"] = st.cases_per_s / (count / t_req)
        res["parity_path"]["speedup_vs_plain"] = sp.cases_per_s / (count / t_req)
        res["plain_device_cell_updates_per_s"] = st.cell_updates / t_dev
        print(f"bench: plain one by one {t_req:.2f} s per stream ({t_dev:.2f} s stepping)", file=sys.stderr,
              flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--path", choices=["fused", "plain", "fast"], default="fast",
                    help="fast: the performance formulation (within 1e-12 of the parity kernels, checked here on the "
                         "benchmark's own run); fused: the parity kernels (the oracle's bits); plain: the per-block "
                         "launches")
    ap.add_argument("--tile", default="auto", help="TIxTJxTK or auto")
    ap.add_argument("--cpu-n", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=96)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe-bw", action="store_true", help="only run the stream-copy probe (1 GiB buffers)")
    ap.add_argument("--serve", type=int, default=16, help="requests of the serving leg (0: skip)")
    ap.add_argument("--lattice", type=int, default=20, help="timed steps of the multi-block lattice leg (0: skip)")
    ap.add_argument("--chem", type=int, default=2000000, help="cells of the chemistry leg (0: skip)")
    ap.add_argument("--chem-cpu", type=int, default=2048, help="cells of the chemistry CPU baseline sample")
    ap.add_argument("--chem-radau", type=int, default=65536, help="cells of the RADAU5 chemistry leg (0: skip)")
    ap.add_argument("--chem-radau-cpu", type=int, default=1024, help="cells of the RADAU5 CPU baseline sample")
    ap.add_argument("--no-serve-plain", action="store_true", help="skip the serving requests one by one on the "
                    "plain per-block path (the comparison path)")
    ap.add_argument("--c1", type=int, default=1, help="the shock-tube leg (1000 cells, 200 steps; 0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU; --gpus {args.gpus} is not supported (the multi-GPU sharding "
              f"of the residual path is a design note, not code)", file=sys.stderr)
        sys.exit(2)

    import mosegpu
    from mosegpu import cases

    if args.probe_bw:
        ctx = mosegpu.Context(cases.air_tables())
        gbs = ctx.bandwidth_probe(1 << 30, 20)
        print(json.dumps({"probe": "stream copy", "bytes_per_launch_read": 1 << 30, "bytes_per_launch_written": 1 << 30,
                          "launches": 21, "GBs": gbs}))
        return

    t_setup = time.perf_counter()
    case = cases.taylor_green(args.n, args.steps)
    ctx = mosegpu.Context(case.tables)
    dc = mosegpu.DeviceCase(ctx, case)
    dc.upload(case.P)
    if args.tile == "auto":
        tile = (0, 0, 0)
    else:
        tile = tuple(int(x) for x in args.tile.split("x"))
    path = {"fused": mosegpu.PATH_FUSED, "plain": mosegpu.PATH_PLAIN, "fast": mosegpu.PATH_FAST}[args.path]
    fused_like = path in (mosegpu.PATH_FUSED, mosegpu.PATH_FAST)
    chosen = dc.tune(path, tile, timing=False)
    t_setup = time.perf_counter() - t_setup
    print(f"bench: case built and tuned in {t_setup:.1f} s, tile {chosen}", file=sys.stderr, flush=True)
    if args.warmup > 0:
        dc.step(args.warmup)
    dc.tune(path, chosen if fused_like else (0, 0, 0), timing=True)
    dc.step(0)                          # drains the stream
    t0 = time.perf_counter()
    rep = dc.step(args.steps)           # returns after the device has finished (the report is read back)
    elapsed = time.perf_counter() - t0
    ms_avg, ms_tot, nlaunch = dc.timing()
    print(f"bench: {args.steps} steps in {elapsed:.3f} s, stage kernel {ms_avg:.3f} ms avg", file=sys.stderr, flush=True)
    cells = case.interior_cells()
    value = cells * args.steps / elapsed
    B = algorithmic_bytes_per_cell_stage(case.nprim, True, False)
    achieved = (B * cells) / (ms_avg * 1e-3) / 1e9 if ms_avg > 0 else None
    traffic, valu = profiled_counts(args.n, case.nprim, True, chosen, args.path) if fused_like else (None, None)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                "traffic": traffic,
                # the share of the launch that FP64 VALU issue alone would take (SQ_INSTS_VALU from the profile)
                "valu_issue_frac": (valu / VALU_ISSUE_PER_S) / (ms_avg * 1e-3) if valu and ms_avg > 0 else None,
                "kernel": ("k_stage (fused residual + RK stage update)" + (", performance formulation"
                                                                              if args.path == "fast" else ""))
                          if fused_like else "face kernels + cell update (one stage)",
                "kernel_ms_avg": ms_avg, "launches": nlaunch, "bytes_per_cell_stage": B}
    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"taylor-green {args.n}^3 periodic, M 0.1, Re 1600, LMRoe(Mco=1) + van Albada, "
                               f"viscous (Sutherland table), SSP-RK3", "cells": cells, "path": args.path,
                   "tile": list(chosen), "parallelism": "single GPU"},
        "roofline": roofline,
        "sim_time": rep.time, "dt_last": rep.dt, "setup_s": t_setup, "device": ctx.info(),
    }
    if args.path == "fast":
        out["parity_check"] = fast_vs_parity(dc, case, chosen, min(args.steps, 20))
    dc.destroy()
    ctx.close()
    if args.lattice > 0:
        out["lattice"] = lattice_leg(args.lattice)
    if args.serve > 0:
        out["serving"] = serving_leg(args.serve, not args.no_serve_plain)
    if args.c1 > 0:
        out["shock_tube"] = shock_tube_leg(0 if args.no_cpu else min(16, os.cpu_count() or 1))
    if args.chem > 0:
        out["chemistry"] = chemistry_leg(args.chem, 0 if args.no_cpu else args.chem_cpu, min(16, os.cpu_count() or 1),
                                         args.chem_radau, 0 if args.no_cpu else args.chem_radau_cpu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps, min(16, os.cpu_count() or 1))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()

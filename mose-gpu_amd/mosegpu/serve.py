"""Continuous-batching serving loop over the C ABI.

Requests arrive as a seeded Poisson stream. Each request is an independent single-block case:
- a 2D four-quadrant Riemann problem at n^2, n log-uniform in [64, 1024]; or
- a 3D viscous Taylor-Green vortex at n^3, n log-uniform in [48, 192].
Each runs 50..500 steps.

The server loop:
1. It admits every prepared request whose device footprint (mose_case_bytes) fits the HBM budget left, oldest
   first.
2. It keeps every resident case stepping on its own stream, a quantum of steps at a time: mose_case_launch queues
   the next quantum as soon as mose_case_poll has collected the last one, so the cases' kernels overlap on the
   device and no case waits for a slower one.
3. It retires finished cases. A retired case gives its state hash, and its memory goes back to the budget.

Requests that arrive while others run are admitted as soon as they are prepared and fit: that is continuous
batching, not a queue in front of a single case.

Host preparation of arrived requests runs on a builder thread, overlapped with the device work of the resident
cases.

Time-to-solution = retirement time - arrival time (host preparation, queueing and device time included).
"""
from __future__ import annotations

import ctypes as C
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import cases
from .core import PATH_FUSED, PATH_PLAIN, Context, DeviceCase, Report, lib


@dataclass
class Request:
    rid: int
    arrival: float          # seconds from the start of the stream
    kind: str               # "riemann2d" | "tgv3d"
    n: int
    steps: int
    seed: int
    done: float = -1.0
    hash: tuple = ()
    admitted: float = -1.0


def request_stream(count: int, rate: float = 20.0, seed: int = 2024, frac_2d: float = 0.5,
                   n2d=(64, 1024), n3d=(48, 192), steps=(50, 500)) -> list[Request]:
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        if rng.uniform() < frac_2d:
            kind, lo, hi = "riemann2d", n2d[0], n2d[1]
        else:
            kind, lo, hi = "tgv3d", n3d[0], n3d[1]
        n = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        out.append(Request(i, t, kind, n, int(rng.integers(steps[0], steps[1] + 1)), int(rng.integers(1 << 30))))
    return out


def build_case(r: Request):
    if r.kind == "riemann2d":
        return cases.riemann2d(r.n, r.steps, seed=r.seed)
    return cases.taylor_green(r.n, r.steps, seed=r.seed)


def case_bytes(case) -> int:
    b = C.c_int64(0)
    lib().mose_case_bytes(C.byref(case.scheme), case.layout.nb,
                          case.layout.dims_arr.ctypes.data_as(C.POINTER(C.c_int)), len(case.records), C.byref(b))
    return int(b.value)


@dataclass
class ServeStats:
    requests: list
    wall: float
    cases_per_s: float
    p50: float
    p99: float
    launches: int           # quanta launched
    max_resident: int
    cell_updates: float
    max_bytes: int = 0
    budget: int = 0
    meta: dict = field(default_factory=dict)


def serve_tile(case, path: int):
    """The stage tile of a served case, by rule instead of the autotuner (which would time candidate tiles on every
    admitted case while the others wait): 8 x 8 x 4 for 3D blocks, 16 x 8 x 1 for 2D ones (a 128-cell tile: a 256-cell
    2D tile's LDS, most of it for the unused one-cell direction, leaves one workgroup per CU; on a 1024^2 Riemann
    case a stage takes 0.25 ms against 0.33 ms with 32 x 8 x 1)."""
    if path == PATH_PLAIN:
        return (0, 0, 0)
    n1, n2, n3 = case.layout.dims[0]
    if n3 == 1:
        return (16, 8, 1) if n2 > 1 else (256, 1, 1)
    return (8, 8, 4)


def serve(reqs: list[Request], budget_bytes: int | None = None, quantum: int = 10, path: int = PATH_FUSED,
          realtime: bool = True, log=None, lookahead: int = 4) -> ServeStats:
    """Run the stream to completion.  realtime: requests become visible at their arrival time (wall clock);
    otherwise all are visible at once (offline throughput).

    A builder thread prepares arrived requests on the host (mesh, metrics, boundary records, initial field), at most
    `lookahead` ahead of admission.  The main thread keeps every resident case busy: each has its own stream, gets its
    next quantum of steps (mose_case_launch) as soon as its last one is collected (mose_case_poll), and retires when
    its steps are done, so no case waits for a slower one and host preparation overlaps device work."""
    ctx_cache = {}

    def ctx_for(case):
        key = case.name in ("riemann2d",) and "euler" or "ns"
        if key not in ctx_cache:
            ctx_cache[key] = Context(case.tables)
        return ctx_cache[key]

    order = sorted(reqs, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    if budget_bytes is None:
        probe = Context(cases.air_tables())
        budget_bytes = int(0.8 * probe.info()["hbm_free"])
        probe.close()
    built: queue.Queue = queue.Queue(maxsize=max(1, lookahead))
    failure = []

    def builder():
        try:
            for r in order:
                if realtime:
                    wait = r.arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(wait)
                case = build_case(r)
                built.put((r, case, case_bytes(case)))
        except BaseException as e:     # handed to the main thread
            failure.append(e)
        finally:
            built.put(None)

    th = threading.Thread(target=builder, name="mose-serve-builder", daemon=True)
    th.start()
    waiting = []        # built, not yet admitted: (req, case, bytes), arrival order
    resident = []       # {"r", "case", "dc", "rem", "nb", "fly"}: fly = steps in flight
    used = 0
    launches = 0
    max_res = 0
    max_used = 0
    cu = 0.0
    stream_done = False
    while True:
        # collect what the builder has ready; block only when there is nothing else to do
        while not stream_done:
            try:
                item = built.get(block=not (resident or waiting))
            except queue.Empty:
                break
            if item is None:
                stream_done = True
            else:
                waiting.append(item)
        if failure:
            raise failure[0]
        # admission: oldest first while the HBM budget allows (a request too large for the budget runs alone)
        while waiting:
            r, case, nb = waiting[0]
            if used + nb > budget_bytes and resident:
                break
            waiting.pop(0)
            ctx = ctx_for(case)
            dc = DeviceCase(ctx, case)
            dc.upload(case.P)
            dc.tune(path, serve_tile(case, path))
            r.admitted = time.perf_counter() - t0
            resident.append({"r": r, "case": case, "dc": dc, "rem": r.steps, "nb": nb, "fly": 0})
            used += nb
        if not resident:
            if stream_done and not waiting:
                break
            continue
        max_res = max(max_res, len(resident))
        max_used = max(max_used, used)
        # every resident case with steps left and nothing in flight gets its next quantum on its own stream
        for it in resident:
            if it["fly"] == 0 and it["rem"] > 0:
                k = min(quantum, it["rem"])
                it["dc"].launch(k)
                it["fly"] = k
                launches += 1
        # collect the quanta that have finished; retire the cases that are done
        progressed = False
        keep = []
        for it in resident:
            if it["fly"] and it["dc"].poll() is not None:
                it["rem"] -= it["fly"]
                cu += it["fly"] * it["case"].interior_cells()
                it["fly"] = 0
                progressed = True
            if it["rem"] <= 0 and it["fly"] == 0:
                r = it["r"]
                r.hash = it["dc"].state_hash()
                r.done = time.perf_counter() - t0
                it["dc"].destroy()
                used -= it["nb"]
                if log:
                    log(f"served {r.kind} n={r.n} steps={r.steps}: {r.done - r.arrival:.3f} s")
            else:
                keep.append(it)
        resident = keep
        if not progressed:
            time.sleep(2e-4)
    th.join()
    wall = time.perf_counter() - t0
    for c in ctx_cache.values():
        c.close()
    lat = np.array([r.done - r.arrival for r in reqs])
    return ServeStats(reqs, wall, len(reqs) / wall, float(np.percentile(lat, 50)), float(np.percentile(lat, 99)),
                      launches, max_res, cu, max_used, budget_bytes)


def run_alone(r: Request, path: int = PATH_FUSED):
    """The same request alone on the device, start to finish.  Returns (state hash, device seconds, request
    seconds): the stepping alone, and the whole request (host preparation, upload, tuning, stepping, hash)."""
    t0 = time.perf_counter()
    case = build_case(r)
    ctx = Context(case.tables)
    dc = DeviceCase(ctx, case)
    dc.upload(case.P)
    dc.tune(path, (0, 0, 0))
    t = time.perf_counter()
    dc.step(r.steps)
    h = dc.state_hash()
    dt = time.perf_counter() - t
    dc.destroy()
    ctx.close()
    return h, dt, time.perf_counter() - t0

"""The serving configuration's cases on the device: the 2D four-quadrant Riemann problem (cases.riemann2d: two
computational directions, MOSE 0 faces in the one-cell direction, extrapolation faces in the plane) against the
oracle, and a served stream that holds cases of the configuration's own sizes (2D n = 512, 3D n = 96) beside small
ones.

The bar is a relative L-infinity <= 1e-12 per primitive variable against the oracle; a served case's final state
must equal the same case run alone bit for bit (field_state_hash), and the stream's smallest case must match the
oracle.
"""
import numpy as np
import pytest

import mosegpu
from mosegpu import serve
from mosegpu import cases
from oracle_bind import OracleCase
from test_gpu_walls_sa import compare, run_device

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("n,steps", [(64, 100), (256, 40)])
@pytest.mark.parametrize("path,tile", [(mosegpu.PATH_FUSED, (0, 0, 0)), (mosegpu.PATH_FUSED, (16, 16, 1)),
                                       (mosegpu.PATH_PLAIN, (0, 0, 0))])
def test_riemann2d_matches_oracle(gpu_ready, n, steps, path, tile):
    case = cases.riemann2d(n, steps, seed=7)
    Pg, rep, chosen = run_device(case, steps, path, tile)
    Po = case.P.copy()
    dts = OracleCase(case, nthreads=16).step(Po, steps)
    assert abs(rep.time - dts.sum()) <= 1e-13 * dts.sum()
    assert compare(case, Pg, Po, f"riemann2d {n}^2 path {path} tile {chosen} x {steps}") <= TOL


def test_served_stream_at_configuration_sizes(gpu_ready):
    reqs = [serve.Request(0, 0.00, "riemann2d", 512, 60, 101),
            serve.Request(1, 0.02, "tgv3d", 96, 50, 102),
            serve.Request(2, 0.03, "riemann2d", 64, 80, 103),
            serve.Request(3, 0.05, "tgv3d", 48, 60, 104),
            serve.Request(4, 0.06, "riemann2d", 181, 50, 105)]
    st = serve.serve(reqs, quantum=10, path=mosegpu.PATH_FUSED, realtime=True)
    assert all(r.done > 0 for r in reqs)
    assert st.max_resident >= 2
    for r in reqs:
        h, _, _ = serve.run_alone(r)
        print(f"served {r.kind} n {r.n} x {r.steps}: hash equal to alone {r.hash == h}", flush=True)
        assert r.hash == h, (r.kind, r.n, r.steps)
    small = min(reqs, key=lambda r: r.n ** (2 if r.kind == "riemann2d" else 3))
    case = serve.build_case(small)
    Po = case.P.copy()
    OracleCase(case, nthreads=16).step(Po, small.steps)
    assert mosegpu.state_hash(case.layout, Po) == small.hash
    print(f"served stream: {st.cases_per_s:.2f} cases/s, p50 {st.p50:.3f} s, p99 {st.p99:.3f} s, max resident "
          f"{st.max_resident}; smallest case ({small.kind} {small.n}) hash equals the oracle's", flush=True)

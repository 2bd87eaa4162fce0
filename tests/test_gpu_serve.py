"""The serving loop: continuous batching of independent cases on one device (mosegpu.serve).

Each served case's final state equals that case run alone, bitwise: the same field_state_hash. Admission respects
the HBM budget, and requests that arrive during the run join the resident set.
"""
import numpy as np
import pytest

from mosegpu import PATH_FUSED, serve

pytestmark = pytest.mark.gpu


def small_stream(count=8, seed=5):
    return serve.request_stream(count, rate=200.0, seed=seed, n2d=(48, 160), n3d=(16, 40), steps=(12, 40))


def test_served_equals_alone(gpu_ready):
    reqs = small_stream()
    st = serve.serve(reqs, quantum=5, path=PATH_FUSED, realtime=True)
    assert all(r.done > 0 for r in reqs)
    assert st.max_resident >= 2          # requests overlapped on the device
    for r in reqs:
        h, _, _ = serve.run_alone(r)
        assert r.hash == h, (r.kind, r.n, r.steps)
    print(f"served {len(reqs)} cases: {st.cases_per_s:.2f} cases/s, p50 {st.p50:.3f} s, p99 {st.p99:.3f} s, "
          f"max resident {st.max_resident}")


def test_budget_limits_residency(gpu_ready):
    reqs = small_stream(6, seed=9)
    sizes = [serve.case_bytes(serve.build_case(r)) for r in reqs]
    budget = int(1.5 * max(sizes))
    st = serve.serve(reqs, budget_bytes=budget, quantum=5, realtime=False)
    assert all(r.done > 0 for r in reqs)
    assert st.max_bytes <= budget                  # admission never exceeds the HBM budget
    assert st.max_resident < len(reqs)             # and the budget did hold requests back


def test_launch_poll_equals_step(gpu_ready):
    """mose_case_launch + mose_case_poll give the state and report of mose_case_step; polling without a launch in
    flight and a second launch before the poll are errors."""
    import time

    from mosegpu import Context, DeviceCase, MoseError, cases
    case = cases.taylor_green(24, 4)
    ctx = Context(case.tables)
    a, b = DeviceCase(ctx, case), DeviceCase(ctx, case)
    a.upload(case.P)
    b.upload(case.P)
    ra = a.step(4)
    b.launch(4)
    with pytest.raises(MoseError):
        b.launch(1)
    with pytest.raises(MoseError):
        b.step(1)                  # a launch awaits its poll
    rb = None
    t0 = time.perf_counter()
    while rb is None and time.perf_counter() - t0 < 60:
        rb = b.poll()
    assert rb is not None
    assert (rb.steps, rb.time, rb.dt) == (ra.steps, ra.time, ra.dt)
    assert a.state_hash() == b.state_hash()
    with pytest.raises(MoseError):
        b.poll()
    a.destroy()
    b.destroy()
    ctx.close()

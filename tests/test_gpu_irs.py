"""Parity of the HIP path with the oracle for the implicit residual smoothing (IRS_apply,
src/lib/numerics/time/explicit/Lib_IRS.f90:53-96; the update on the prescaled residual, newstate_cons_body.inc:47).

The smoothing needs the whole residual, so it runs on the plain path: the residual of every cell into an array, its
scaling, per direction the residual's ghost layer and the Jacobi sweeps of each block, then the update.  Both sides
run the same statements: the cases without far-field faces or the SA source must equal the oracle bit for bit, the SA
channel is held to 1e-12 (its real powers come from the device's and the host's pow)."""
import numpy as np
import pytest

import mosegpu
from mosegpu import Context, DeviceCase, MoseError, cases
from test_gpu_prim import moving_box, parity, prim
from test_gpu_walls_sa import interiors, run_device

pytestmark = pytest.mark.gpu

PLAIN = mosegpu.PATH_PLAIN


def irs(case, beta=0.5, J=2):
    case.scheme.irs = 1
    case.scheme.irs_beta = beta
    case.scheme.irs_njacobi = J
    return case


def test_irs_sod(gpu_ready):
    parity(irs(cases.sod(400, 60), 0.6, 2), 60, PLAIN, (0, 0, 0), "Sod 400 with IRS")


@pytest.mark.parametrize("integration", ["cons", "prim"])
def test_irs_taylor_green(gpu_ready, integration):
    c = irs(cases.taylor_green(16, 8), 0.5, 3)
    if integration == "prim":
        prim(c)
    parity(c, 8, PLAIN, (0, 0, 0), f"Taylor-Green 16^3 with IRS ({integration})")


def test_irs_four_species(gpu_ready):
    parity(irs(moving_box(4), 0.4, 2), 6, PLAIN, (0, 0, 0), "moving viscous box, 4 species, with IRS (prim)")


def test_irs_channel_sa(gpu_ready):
    c = irs(cases.channel_lattice(nb=(2, 2, 2), wmin=4, wmax=9, seed=11, nsteps=5, riemann="LMRoe"), 0.5, 2)
    parity(c, 5, PLAIN, (0, 0, 0), "SA channel 2x2x2 with IRS", bitwise=False)


def test_irs_needs_the_plain_path_and_changes_the_result(gpu_ready):
    a = cases.taylor_green(16, 4)
    b = irs(cases.taylor_green(16, 4), 0.5, 2)
    ctx = Context(b.tables)
    dc = DeviceCase(ctx, b)
    with pytest.raises(MoseError):
        dc.tune(mosegpu.PATH_FUSED, (8, 8, 4))
    dc.destroy(); ctx.close()
    Pa, _, _ = run_device(a, 4, PLAIN, (0, 0, 0))
    Pb, _, _ = run_device(b, 4, PLAIN, (0, 0, 0))
    d = np.abs(interiors(a, Pa) - interiors(b, Pb)).max() / np.abs(interiors(a, Pa)).max()
    print(f"Taylor-Green 16^3 x 4, with and without IRS: relative difference {d:.3e}")
    assert d > 1e-9

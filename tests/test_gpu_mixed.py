"""Mixed-precision path (FP32 stencil arithmetic, FP64 state / pressure
recovery / RK accumulation / CFL): error against the FP64 path, which is
bit-identical to the CPU oracle.  Tolerances are relative max-norms per
conserved field, normalised by the field's largest magnitude (rho v and
rho w by the momentum scale rho u)."""
import numpy as np
import pytest

import freeshearlayer_amd as fsl

pytestmark = pytest.mark.gpu


def run(case, U, u_in, precision, sgs, steps, cfl=0.5):
    cfg = fsl.config(sgs_model=sgs, precision=precision, fused=1, cfl_mode=fsl.FSL_CFL_ACOUSTIC, cfl=cfl)
    with fsl.Context(case, cfg) as ctx:
        ctx.set_inflow(u_in)
        ctx.upload_state(U)
        ctx.set_lcg(0.317391239817, 0)
        dts = []
        for _ in range(steps):
            ctx.step()
            dts.append(ctx.dt)
        return ctx.download_state()[:, 1:-1, 1:-1, 1:-1], np.array(dts), ctx.conservation_sums()


def rel_err(a, b):
    scale = [np.abs(b[0]).max(), np.abs(b[1]).max(), np.abs(b[1]).max(), np.abs(b[1]).max(),
             np.abs(b[4]).max()]
    return [float(np.abs(a[f] - b[f]).max() / scale[f]) for f in range(5)]


@pytest.mark.parametrize("sgs", [fsl.FSL_SGS_CONSTANT, fsl.FSL_SGS_DYNAMIC])
def test_mixed_vs_fp64_one_and_twenty_steps(sgs):
    case = fsl.case_reference(64, 48, 32)
    U, u_in, mu = fsl.ic_shear_layer(case)
    case.mu_L = mu
    for steps, tol in ((1, 2e-6), (20, 2e-4)):
        a, dta, sa = run(case, U, u_in, fsl.FSL_PREC_MIXED, sgs, steps)
        b, dtb, sb = run(case, U, u_in, fsl.FSL_PREC_FP64, sgs, steps)
        e = rel_err(a, b)
        print("mixed vs fp64, sgs=%d, %d steps: %s" % (sgs, steps, ["%.2e" % x for x in e]))
        assert max(e) < tol, e
        assert np.all(np.isfinite(a))
        # dt from the FP64 CFL reduction of slightly different states
        assert abs(dta[-1] - dtb[-1]) / dtb[-1] < 1e-5
        # mass and energy totals agree to FP64-accumulation precision of the flux differences
        assert abs(sa[0] - sb[0]) / abs(sb[0]) < 1e-9
        assert abs(sa[4] - sb[4]) / abs(sb[4]) < 1e-9


def test_mixed_reference_case_runs_and_stays_close():
    case = fsl.case_reference(48, 40, 16)
    U = fsl.ic_reference(case)
    u_in = np.where(np.arange(1, 41) <= 20, 76.4, 200.0)
    a, _, _ = run(case, U, u_in, fsl.FSL_PREC_MIXED, fsl.FSL_SGS_DYNAMIC, 10)
    b, _, _ = run(case, U, u_in, fsl.FSL_PREC_FP64, fsl.FSL_SGS_DYNAMIC, 10)
    e = rel_err(a, b)
    print("reference case, 10 steps:", ["%.2e" % x for x in e])
    assert max(e) < 1e-3

"""The time-step guards on the GPU.

* The acoustic rule (every GPU configuration's time step):
  dt = cfl / max over interior cells of (|u|+c)/dx + (|v|+c)/dy + (|w|+c)/dz
  with c = sqrt(gamma p / rho), p = (gamma-1)(E - rho|u|^2/2), checked against
  a numpy evaluation of that formula (the reference has only the convective
  rule of src/helpers.c:23-52, which the oracle pins bit for bit elsewhere).
* Non-finite states: one NaN cell anywhere makes the CFL maximum NaN (the
  reference's comparison would skip it, src/helpers.c:39-41), so the step
  fails with FSL_ERR_NONFINITE and the context keeps its last valid dt; on
  the fused path the maximum is produced by stage 3 of the previous step."""
import numpy as np
import pytest

import freeshearlayer_amd as fsl
from fsl_test_util import random_state

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def acoustic_max(U, dx, dy, dz):
    q = U[:, 1:-1, 1:-1, 1:-1]
    rho = q[0]
    u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
    p = (q[4] - 0.5 * rho * (u * u + v * v + w * w)) * (GAMMA - 1)
    c = np.sqrt(GAMMA * p / rho)
    return float(np.max((np.abs(u) + c) / dx + (np.abs(v) + c) / dy + (np.abs(w) + c) / dz))


@pytest.mark.parametrize("precision", [fsl.FSL_PREC_FP64, fsl.FSL_PREC_MIXED])
@pytest.mark.parametrize("mesh", [(16, 12, 8), (37, 21, 19), (64, 48, 32)])
def test_acoustic_cfl_matches_numpy(mesh, precision):
    L, H, D = mesh
    case = fsl.case_reference(L, H, D)
    cfg = fsl.config(sgs_model=fsl.FSL_SGS_CONSTANT, cfl_mode=fsl.FSL_CFL_ACOUSTIC, cfl=0.5, precision=precision,
                     fused=1)
    U = random_state(L, H, D, seed=11)
    # a sign mix in every velocity component so |u|, |v|, |w| matter
    U[1] *= np.where(np.arange(D + 2) % 3 == 0, -1.0, 1.0)[None, None, :]
    with fsl.Context(case, cfg) as ctx:
        ctx.upload_state(U)
        m = ctx.check_co_num()
        ref = acoustic_max(U, case.deltaX, case.deltaY, case.deltaZ)
        # FP64 throughout; the device uses the same operations in another order
        assert abs(m - ref) <= 1e-14 * ref, (m, ref)
        assert abs(ctx.dt - 0.5 / m) <= 1e-15 * ctx.dt


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("mode", [fsl.FSL_CFL_CONVECTIVE, fsl.FSL_CFL_ACOUSTIC])
def test_one_nan_cell_fails_the_step_and_keeps_dt(fused, mode):
    L, H, D = 16, 12, 8
    case = fsl.case_reference(L, H, D)
    cfg = fsl.config(sgs_model=fsl.FSL_SGS_CONSTANT, cfl_mode=mode, cfl=0.5, fused=fused)
    U = random_state(L, H, D, seed=3)
    with fsl.Context(case, cfg) as ctx:
        ctx.upload_state(U)
        ctx.set_lcg(0.317391239817, 0)
        ctx.step()
        dt_ok = ctx.dt
        V = ctx.download_state()
        V[0, 9, 5, 3] = np.nan          # one interior cell's density
        ctx.upload_state(V)
        with pytest.raises(fsl.FslError, match=r"\(%d\)" % fsl.FSL_ERR_NONFINITE):
            ctx.step()
        assert ctx.dt == dt_ok


@pytest.mark.parametrize("precision", [fsl.FSL_PREC_FP64, fsl.FSL_PREC_MIXED])
def test_nan_made_inside_a_step_fails_the_next_rescale(precision):
    """On the fused path stage 3 reduces the CFL quantity of the state it
    writes, and the next step's rescale uses that maximum.  A NaN in the
    mu_SGS ghost layer (which the boundary faces sample, src/fluxes.c:191;
    the dynamic model writes the interior only) turns the x-inflow cells NaN
    during the step: the step itself passes its rescale on the clean input,
    the next one must fail."""
    L, H, D = 16, 12, 8
    case = fsl.case_reference(L, H, D)
    cfg = fsl.config(sgs_model=fsl.FSL_SGS_DYNAMIC, cfl_mode=fsl.FSL_CFL_ACOUSTIC, cfl=0.5, fused=1,
                     precision=precision)
    U = random_state(L, H, D, seed=5)
    with fsl.Context(case, cfg) as ctx:
        ctx.upload_state(U)
        ctx.set_lcg(0.317391239817, 0)
        mu = np.zeros((L + 2, H + 2, D + 2))
        mu[0, 6, 4] = np.nan
        ctx.upload_mu(mu)
        ctx.step()
        dt1 = ctx.dt
        assert np.isnan(ctx.download_state()[:, 1, 6, 4]).any()
        with pytest.raises(fsl.FslError, match=r"\(%d\)" % fsl.FSL_ERR_NONFINITE):
            ctx.step()
        assert ctx.dt == dt1

"""Averaged dynamic coefficient (opt-in; the reference's model is pointwise):
the device's z-line and x-z-plane averages against numpy averages of the
oracle's Germano terms L:M and M:M (tests/oracle_lib.py dynamic_terms, the
intermediates of the oracle's pointwise model, which is pinned bit for bit
to the compiled reference).  Cd = -1/2 sum L:M / (sum M:M + 1e-30) clipped
to [0, 0.15], mu_SGS = Cd rho Delta^2 |S|.

Tolerances: FP64 contexts sum in another order than numpy (rtol 1e-10 on
mu_SGS); the FP32 model of mixed contexts forms the terms in FP32 (atol 2e-4
of the largest mu_SGS).  Switching back to pointwise reproduces the oracle bit
for bit."""
import numpy as np
import pytest

import freeshearlayer_amd as fsl
from fsl_test_util import assert_bitwise, make_pair

pytestmark = pytest.mark.gpu


def averaged_mu(lm, mm, fac, mode):
    I = (slice(1, -1),) * 3
    a, b = lm[I], mm[I]
    if mode == fsl.FSL_SGS_AVG_Z:
        sa, sb = a.sum(2, keepdims=True), b.sum(2, keepdims=True)
    else:
        sa, sb = a.sum((0, 2), keepdims=True), b.sum((0, 2), keepdims=True)
    Cd = np.clip(-0.5 * (sa / (sb + 1e-30)), 0.0, 0.15)
    mu = np.zeros_like(fac)
    mu[I] = Cd * fac[I]
    return mu, Cd


@pytest.mark.parametrize("mode", [fsl.FSL_SGS_AVG_Z, fsl.FSL_SGS_AVG_XZ])
@pytest.mark.parametrize("precision,fp32_form", [(fsl.FSL_PREC_FP64, False), (fsl.FSL_PREC_MIXED, False),
                                                (fsl.FSL_PREC_MIXED, True)])
def test_averaged_coefficient_matches_oracle_terms(mode, precision, fp32_form, monkeypatch):
    # a mixed context runs the FP64 model unless FSL_SGS_FP32=1 selects the FP32 form
    if fp32_form:
        monkeypatch.setenv("FSL_SGS_FP32", "1")
    else:
        monkeypatch.delenv("FSL_SGS_FP32", raising=False)
    exact = not fp32_form
    L, H, D = 20, 18, 24          # tiles that do not divide the mesh
    orc, ctx = make_pair(L, H, D, dynamic=1, seed=31, precision=precision)
    with ctx:
        ctx.ghost_fill()
        orc.ghost()
        lm, mm, fac = orc.dynamic_terms()
        ref, Cd = averaged_mu(lm, mm, fac, mode)
        assert 0 < np.count_nonzero(Cd) and np.count_nonzero(Cd < 0.15) > 0
        ctx.set_sgs_average(mode)
        assert ctx.sgs_average() == mode
        ctx.dynamic_smagorinsky()
        mu = ctx.download_mu()
        err = np.abs(mu - ref).max() / np.abs(ref).max()
        print("mode %d precision %d: max |mu - ref| / max mu = %.2e" % (mode, precision, err))
        if exact:
            np.testing.assert_allclose(mu, ref, rtol=1e-10, atol=1e-14 * np.abs(ref).max())
        else:
            assert err < 2e-4
        # the averaged Cd is constant along each z line (and, for XZ, over each plane)
        I = (slice(1, -1),) * 3
        c = np.divide(mu[I], fac[I], out=np.zeros_like(mu[I]), where=fac[I] > 0)
        spread = np.ptp(c, axis=2 if mode == fsl.FSL_SGS_AVG_Z else (0, 2))
        assert spread.max() < (1e-12 if exact else 1e-5)
        # back to the reference's pointwise model
        ctx.set_sgs_average(fsl.FSL_SGS_AVG_NONE)
        ctx.dynamic_smagorinsky()
        orc.dynamic()
        if exact:
            assert_bitwise(orc.cell(10), ctx.download_mu(), "pointwise mu_SGS after averaging")


def test_averaged_step_runs_and_rejects_bad_modes():
    orc, ctx = make_pair(24, 20, 16, dynamic=1, seed=32, fused=1)
    with ctx:
        with pytest.raises(fsl.FslError):
            ctx.set_sgs_average(3)
        ctx.set_sgs_average(fsl.FSL_SGS_AVG_XZ)
        for _ in range(3):
            ctx.step()
        U = ctx.download_state()
        assert np.all(np.isfinite(U))

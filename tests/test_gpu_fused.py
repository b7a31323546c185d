"""GPU parity of the fused stage path: one kernel per stage (plus the ghost
fill and, once per step, the dynamic model) against the CPU oracle, bit for
bit, on meshes that do and do not divide the 14 x 14 workgroup tile, and
against the unfused device path at the benchmark size."""
import numpy as np
import pytest

import freeshearlayer_amd as fsl
from fsl_test_util import assert_bitwise, make_pair, random_state

pytestmark = pytest.mark.gpu

MESHES = [(16, 12, 8), (5, 10, 3), (30, 29, 15), (9, 44, 40), (40, 10, 1)]


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("dynamic", [0, 1])
def test_fused_steps_match_oracle(mesh, dynamic):
    orc, ctx = make_pair(*mesh, dynamic=dynamic, seed=31, fused=1)
    for n in range(3):
        ctx.step()
        orc.step()
        assert ctx.dt == orc.get("dt"), "dt at step %d" % (n + 1)
        assert_bitwise(orc.state()[:, 1:-1, 1:-1, 1:-1], ctx.download_state()[:, 1:-1, 1:-1, 1:-1],
                       "interior U after step %d" % (n + 1))


@pytest.mark.parametrize("dynamic", [0, 1])
def test_fused_reference_case(dynamic):
    orc, ctx = make_pair(16, 12, 8, dynamic=dynamic, fused=1)
    for n in range(8):
        ctx.step()
        orc.step()
        assert_bitwise(orc.state()[:, 1:-1, 1:-1, 1:-1], ctx.download_state()[:, 1:-1, 1:-1, 1:-1],
                       "interior U after step %d" % (n + 1))


@pytest.mark.parametrize("quirk", [0, 1])
def test_fused_equals_unfused_at_bench_size(quirk):
    """Size-independent property at 192 x 128 x 96: the fused and unfused
    device paths give the same bits after two steps."""
    L, H, D = 192, 128, 96
    case = fsl.case_reference(L, H, D)
    U, u_in, mu = fsl.ic_shear_layer(case)
    case.mu_L = mu
    out = []
    for fused in (0, 1):
        cfg = fsl.config(sgs_model=fsl.FSL_SGS_DYNAMIC, fused=fused, z_quirk=quirk,
                         cfl_mode=fsl.FSL_CFL_ACOUSTIC, cfl=0.5)
        with fsl.Context(case, cfg) as ctx:
            ctx.set_inflow(u_in)
            ctx.upload_state(U)
            ctx.set_lcg(0.317391239817, 0)
            ctx.run(2)
            out.append(ctx.download_state()[:, 1:-1, 1:-1, 1:-1])
    assert_bitwise(out[0], out[1], "fused vs unfused")
    assert np.all(np.isfinite(out[1]))


@pytest.mark.parametrize("dynamic", [0, 1])
def test_hundred_steps_shear_layer_divergence(dynamic):
    """The CPU-runnable configuration: the shear layer (tanh profile, Mc 0.3,
    U2/U1 0.5, Re 500, 2 % seeded perturbations) on 64 x 48 x 32 cells, 3-stage
    TVD RK, the reference's time-step rule, 100 steps on the fused FP64 path
    against the oracle.  The
    divergence (max relative difference per conserved field) is reported at
    every step and must stay 0: the path is bit-identical, so no drift can
    accumulate."""
    from oracle_lib import Oracle
    from freeshearlayer_amd import shear_case
    L, H, D = 64, 48, 32
    case, U, u_in, shear = shear_case.initial(L, H, D)
    # the reference's convective rescale with its own maxCoNum 0.1: an acoustic
    # CFL number of about 0.4 on this case (convective 0.5 would be about 2)
    assert case.maxCoNum == 0.1
    orc = Oracle(L, H, D, dynamic=dynamic)
    orc.set("mu_L", case.mu_L)
    orc.set("BL_HIG", case.BL_HIG)
    orc.u_in()[:] = u_in
    orc.derive()
    orc.set_state(U)
    orc.set("X", 0.317391239817)
    cfg = fsl.config(sgs_model=fsl.FSL_SGS_DYNAMIC if dynamic else fsl.FSL_SGS_CONSTANT, fused=1)
    worst = 0.0
    with fsl.Context(case, cfg) as ctx:
        ctx.set_inflow(u_in)
        ctx.upload_state(U)
        ctx.set_lcg(0.317391239817, 0)
        for n in range(1, 101):
            ctx.step()
            orc.step()
            a, b = ctx.download_state()[:, 1:-1, 1:-1, 1:-1], orc.state()[:, 1:-1, 1:-1, 1:-1]
            div = [float(np.abs(a[f] - b[f]).max() / np.abs(b[f]).max()) for f in range(5)]
            worst = max(worst, max(div))
            if n in (1, 10, 50, 100):
                print("step %3d: dt %.6e, divergence per field %s" % (n, ctx.dt, " ".join("%.1e" % d for d in div)))
            assert ctx.dt == orc.get("dt"), "dt at step %d" % n
            assert max(div) == 0.0, "step %d: %s" % (n, div)
    orc.close()
    assert worst == 0.0


def test_fused_matches_oracle_at_192x128x96():
    """The GPU configuration's mesh: the shear layer on 192 x 128 x 96 cells
    with dynamic Smagorinsky, two steps of the fused FP64 path against the
    CPU oracle, bit for bit (state and dt; the oracle needs a few seconds per
    step here)."""
    from oracle_lib import Oracle
    from freeshearlayer_amd import shear_case
    L, H, D = 192, 128, 96
    case, U, u_in, shear = shear_case.initial(L, H, D)
    orc = Oracle(L, H, D, dynamic=1)
    orc.set("mu_L", case.mu_L)
    orc.set("BL_HIG", case.BL_HIG)
    orc.u_in()[:] = u_in
    orc.derive()
    orc.set_state(U)
    orc.set("X", 0.317391239817)
    cfg = fsl.config(sgs_model=fsl.FSL_SGS_DYNAMIC, fused=1)
    with fsl.Context(case, cfg) as ctx:
        ctx.set_inflow(u_in)
        ctx.upload_state(U)
        ctx.set_lcg(0.317391239817, 0)
        for n in range(1, 3):
            ctx.step()
            orc.step()
            assert ctx.dt == orc.get("dt"), "dt at step %d" % n
            assert_bitwise(orc.state()[:, 1:-1, 1:-1, 1:-1], ctx.download_state()[:, 1:-1, 1:-1, 1:-1],
                           "interior U after step %d" % n)
    orc.close()

"""GPU parity of the fused stage path: one kernel per stage (plus the ghost
fill and, once per step, the dynamic model) against the CPU oracle, bit for
bit, on meshes that do and do not divide the 14 x 14 workgroup tile, and
against the unfused device path at the benchmark size; the fast build of the
FP64 stage kernel (FSL_PREC_FP64_FAST) within 1e-12 of the oracle after one
step, and its divergence over 100 steps."""
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


def _shear_oracle(L, H, D, dynamic):
    from oracle_lib import Oracle
    from freeshearlayer_amd import shear_case
    case, U, u_in, shear = shear_case.initial(L, H, D)
    orc = Oracle(L, H, D, dynamic=dynamic)
    orc.set("mu_L", case.mu_L)
    orc.set("BL_HIG", case.BL_HIG)
    orc.u_in()[:] = u_in
    orc.derive()
    orc.set_state(U)
    orc.set("X", 0.317391239817)
    return case, U, u_in, orc


def _rel_divergence(a, b):
    """max |a - b| / max |b| per conserved field (interior cells)"""
    return [float(np.abs(a[f] - b[f]).max() / np.abs(b[f]).max()) for f in range(5)]


@pytest.mark.parametrize("mesh,dynamic", [((64, 48, 32), 0), ((192, 128, 96), 1)])
def test_fast_fp64_one_step_within_1e12_of_oracle(mesh, dynamic):
    """The fast build of the FP64 stage kernel (FSL_PREC_FP64_FAST: FMA
    contraction, quotients by the refined reciprocal) against the CPU oracle
    after one step on the shear layer: 64 x 48 x 32 with constant Cs and
    192 x 128 x 96 with the dynamic model.  Tolerance: 1e-12 relative in the
    max norm of each conserved field.  The time step of the first step comes
    from the initial state (the same kernel in both builds) and is equal."""
    L, H, D = mesh
    case, U, u_in, orc = _shear_oracle(L, H, D, dynamic)
    cfg = fsl.config(sgs_model=fsl.FSL_SGS_DYNAMIC if dynamic else fsl.FSL_SGS_CONSTANT, fused=1,
                     precision=fsl.FSL_PREC_FP64_FAST)
    with fsl.Context(case, cfg) as ctx:
        ctx.set_inflow(u_in)
        ctx.upload_state(U)
        ctx.set_lcg(0.317391239817, 0)
        ctx.step()
        orc.step()
        assert ctx.dt == orc.get("dt")
        a, b = ctx.download_state()[:, 1:-1, 1:-1, 1:-1], orc.state()[:, 1:-1, 1:-1, 1:-1]
    orc.close()
    div = _rel_divergence(a, b)
    print("fast FP64, one step, %dx%dx%d: divergence per field %s" % (L, H, D, " ".join("%.1e" % d for d in div)))
    assert np.all(np.isfinite(a))
    assert max(div) <= 1e-12, div
    assert max(div) > 0.0, "the fast build is expected to round differently somewhere"


def test_fast_fp64_hundred_step_divergence():
    """Divergence of the fast FP64 build from the parity path (itself equal
    to the oracle bit for bit over these 100 steps) on the 64 x 48 x 32 shear
    layer with the dynamic model, reported at steps 1, 10, 50 and 100; it
    grows from rounding level and stays below 1e-10 relative here."""
    from freeshearlayer_amd import shear_case
    L, H, D = 64, 48, 32
    case, U, u_in, shear = shear_case.initial(L, H, D)
    ctxs = []
    for prec in (fsl.FSL_PREC_FP64, fsl.FSL_PREC_FP64_FAST):
        ctx = fsl.Context(case, fsl.config(sgs_model=fsl.FSL_SGS_DYNAMIC, fused=1, precision=prec))
        ctx.set_inflow(u_in)
        ctx.upload_state(U)
        ctx.set_lcg(0.317391239817, 0)
        ctxs.append(ctx)
    worst = []
    try:
        for n in range(1, 101):
            for ctx in ctxs:
                ctx.step()
            if n in (1, 10, 50, 100):
                b, a = (c.download_state()[:, 1:-1, 1:-1, 1:-1] for c in ctxs)
                div = _rel_divergence(a, b)
                worst.append(max(div))
                print("step %3d: dt %.9e / %.9e, divergence per field %s" % (
                    n, ctxs[0].dt, ctxs[1].dt, " ".join("%.1e" % d for d in div)))
                assert np.all(np.isfinite(a))
                assert abs(ctxs[1].dt - ctxs[0].dt) <= 1e-10 * ctxs[0].dt
    finally:
        for ctx in ctxs:
            ctx.close()
    assert worst[0] <= 1e-12 and worst[-1] <= 1e-10, worst


@pytest.mark.parametrize("env", [{"FSL_FP64_FIELD_RSRC": "1"}, {"FSL_FUSED_VARIANT": "1"}],
                         ids=["field-resources", "one-wave-per-simd"])
@pytest.mark.parametrize("dynamic", [0, 1])
def test_other_stage_kernel_forms_match_oracle(monkeypatch, env, dynamic):
    """The FP64 stage kernel's other compiled forms against the oracle, bit
    for bit: one buffer resource per field and plane (the form of meshes
    beyond (4 G + one plane) * 8 >= 2^32 bytes, which FSL_FP64_FIELD_RSRC=1
    forces at any size) and the one-wave-per-SIMD variant
    (FSL_FUSED_VARIANT=1)."""
    for key, val in env.items():
        monkeypatch.setenv(key, val)
    orc, ctx = make_pair(30, 29, 15, dynamic=dynamic, seed=31, fused=1)
    for n in range(3):
        ctx.step()
        orc.step()
        assert ctx.dt == orc.get("dt"), "dt at step %d" % (n + 1)
        assert_bitwise(orc.state()[:, 1:-1, 1:-1, 1:-1], ctx.download_state()[:, 1:-1, 1:-1, 1:-1],
                       "interior U after step %d" % (n + 1))


def test_fast_build_field_resources_match_its_plane_resources(monkeypatch):
    """The fast FP64 build in its per-field-resource form gives the bits of
    its one-resource-per-plane form (the same arithmetic, other addressing)."""
    out = []
    for forced in ("0", "1"):
        monkeypatch.setenv("FSL_FP64_FIELD_RSRC", forced)
        orc, ctx = make_pair(30, 29, 15, dynamic=1, seed=31, fused=1, precision=fsl.FSL_PREC_FP64_FAST)
        orc.close()
        ctx.run(3)
        out.append(ctx.download_state()[:, 1:-1, 1:-1, 1:-1])
        ctx.close()
    assert_bitwise(out[0], out[1], "fast build, per-field vs per-plane resources")

"""GPU parity of the unfused path: each reference entry point on the device
against the CPU oracle on identical seeded inputs.  FP64, no contraction,
same operation order, so the bar is bit-for-bit equality."""
import numpy as np
import pytest

import freeshearlayer_amd as fsl
from fsl_test_util import assert_bitwise, make_pair, random_state

pytestmark = pytest.mark.gpu

MESHES = [(16, 12, 8), (7, 6, 5), (24, 10, 33)]


@pytest.mark.parametrize("mesh", MESHES)
def test_ghost_fill(mesh):
    orc, ctx = make_pair(*mesh, dynamic=1, seed=11)
    ctx.ghost_fill()
    orc.ghost()
    assert_bitwise(orc.state(), ctx.download_state(), "ghost-filled state")


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("quirk", [1, 0])
def test_reconstruction(mesh, quirk):
    orc, ctx = make_pair(*mesh, dynamic=1, seed=12, z_quirk=quirk)
    orc.set("z_quirk", quirk)
    ctx.ghost_fill(); orc.ghost()
    ctx.reconstruct(); orc.recon()
    for fid in range(30):
        assert_bitwise(orc.face(fid), ctx.download_face(fid), "face array %d" % fid)


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("block", [(2, 6, 12.5, -3.0, 7.25), (6, 2, -9.0, 4.0, -1.5), (3, 3, 5.0, 5.0, 5.0)])
def test_interface_bc(mesh, block):
    orc, ctx = make_pair(*mesh, dynamic=1, seed=13)
    kmin, kmax, ud, vd, wd = block
    kmin, kmax = min(kmin, mesh[2]), min(kmax, mesh[2])
    for k, v in zip(("k_min", "k_max", "Ud", "Vd", "Wd"), (kmin, kmax, ud, vd, wd)):
        orc.set(k, v)
    ctx.set_block(kmin, kmax, ud, vd, wd)
    ctx.ghost_fill(); orc.ghost()
    ctx.reconstruct(); orc.recon()
    ctx.iface_bc(); orc.iface()
    for fid in range(30):
        assert_bitwise(orc.face(fid), ctx.download_face(fid), "face array %d" % fid)


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("dynamic", [0, 1])
def test_fluxes(mesh, dynamic):
    orc, ctx = make_pair(*mesh, dynamic=dynamic, seed=14)
    orc.set("k_min", 1); orc.set("k_max", 3); orc.set("Ud", 10.0)
    ctx.set_block(1, 3, 10.0, 0.0, 0.0)
    ctx.ghost_fill(); orc.ghost()
    ctx.reconstruct(); orc.recon()
    ctx.iface_bc(); orc.iface()
    ctx.fluxes(); orc.fluxes()
    if dynamic:
        assert_bitwise(orc.cell(10), ctx.download_mu(), "mu_SGS")
    for fid in (0, 1, 2, 3, 4, 10, 11, 12, 13, 14, 20, 21, 22, 23, 24):
        assert_bitwise(orc.face(fid), ctx.download_face(fid), "flux array %d" % fid)


@pytest.mark.parametrize("mesh", [(16, 12, 8), (9, 8, 6)])
def test_dynamic_smagorinsky_and_filter(mesh):
    orc, ctx = make_pair(*mesh, dynamic=1, seed=15)
    ctx.ghost_fill(); orc.ghost()
    ctx.dynamic_smagorinsky(); orc.dynamic()
    assert_bitwise(orc.cell(10), ctx.download_mu(), "mu_SGS")
    a = random_state(*mesh, seed=16)[1]
    assert_bitwise(orc.filter(a), ctx.filter(a), "filter_")


@pytest.mark.parametrize("nstages_stage", [(3, 1), (3, 2), (3, 3), (2, 1), (2, 2)])
def test_evolution(nstages_stage):
    """Evolution(numStages, Stage) on injected fluxes and stage states; with
    numStages = 2 the two-stage update falls through into the three-stage one
    (src/evolution.c:19-22)."""
    nst, stage = nstages_stage
    mesh = (16, 12, 8)
    orc, ctx = make_pair(*mesh, dynamic=0, seed=17)
    Up = random_state(*mesh, seed=18)
    for f in range(5):
        orc.cell(5 + f)[...] = Up[f]
    ctx.upload_state(Up, which=1)
    rng = np.random.default_rng(19)
    for fid in (0, 1, 2, 3, 4, 10, 11, 12, 13, 14, 20, 21, 22, 23, 24):
        a = rng.uniform(-1e3, 1e3, orc.face(fid).shape)
        orc.face(fid)[...] = a
        ctx.upload_face(fid, a)
    dt = 1.3e-4
    ctx.set_dt(dt)
    orc.set("dt", dt)
    for k, d in (("dtX", 1.6666667), ("dtY", 1.0), ("dtZ", 1.5)):
        orc.set(k, dt / d)
    orc.evolution(nst, stage)
    ctx.evolution(nst, stage)
    assert_bitwise(orc.state(), ctx.download_state(0), "U")
    Upo = np.stack([orc.cell(5 + f) for f in range(5)])
    assert_bitwise(Upo, ctx.download_state(1), "Up")
    assert ctx.cur_is_p == orc.cur_is_p()


@pytest.mark.parametrize("dynamic", [0, 1])
def test_check_co_num(dynamic):
    orc, ctx = make_pair(16, 12, 8, dynamic=dynamic, seed=20)
    m_gpu = ctx.check_co_num()
    m_cpu = orc.check_co_num()
    assert m_gpu == m_cpu
    assert ctx.dt == orc.get("dt")


@pytest.mark.parametrize("dynamic", [0, 1])
@pytest.mark.parametrize("nstages", [3, 2])
def test_full_steps_reference_case(dynamic, nstages):
    """Reference two-stream case, driver loop (CFL, block draw, stages) on the
    device vs the oracle's step, bit-for-bit after every step."""
    orc, ctx = make_pair(16, 12, 8, dynamic=dynamic, nstages=nstages)
    for n in range(6):
        ctx.step()
        orc.step()
        assert ctx.dt == orc.get("dt"), "dt at step %d" % (n + 1)
        assert_bitwise(orc.state(), ctx.download_state(0), "U after step %d" % (n + 1))
    lcg = ctx.lcg()
    assert lcg["X"] == orc.get("X") and lcg["k_min"] == orc.get("k_min")


@pytest.mark.parametrize("dynamic", [0, 1])
def test_full_steps_random_state(dynamic):
    orc, ctx = make_pair(20, 14, 10, dynamic=dynamic, seed=21)
    for n in range(3):
        ctx.step()
        orc.step()
        assert_bitwise(orc.state(), ctx.download_state(0), "U after step %d" % (n + 1))

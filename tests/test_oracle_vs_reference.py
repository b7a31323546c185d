"""The CPU oracle against the compiled reference (oracle/_ref, built from the
reference's own src/ files by oracle/Makefile), bit for bit, on meshes and
states beyond the committed fixtures.  Skipped where the reference library
has not been built (it is built only where the reference sources exist)."""
import numpy as np
import pytest

from fsl_test_util import assert_bitwise, random_state
from oracle_lib import Oracle, RefLib, ref_available

X0 = 0.317391239817
VARIANTS = [v for v in ("cs", "dyn")]


def need(variant):
    if not ref_available(variant):
        pytest.skip("compiled reference not built (oracle/_ref)")


def pair(variant, L, H, D, seed=None, nstages=3):
    need(variant)
    ref = RefLib(variant)
    bl = min(5, H - H // 2)
    ref.initialize(L, H, D, nStages=nstages, BL_HIG=bl)
    ref.set_time(0.0007, X0, 0)
    orc = Oracle(L, H, D, dynamic=int(variant == "dyn"))
    orc.set("nStages", nstages)
    orc.set("BL_HIG", bl)
    if seed is None:
        orc.ic_reference()
    else:
        U = random_state(L, H, D, seed)
        orc.set_state(U)
        ref.set_state(U)
    return ref, orc


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("mesh", [(5, 4, 3), (13, 10, 7), (8, 16, 12)])
@pytest.mark.parametrize("seed", [None, 3])
def test_steps_bitwise(variant, mesh, seed):
    ref, orc = pair(variant, *mesh, seed=seed)
    for n in range(4):
        ref.step()
        orc.step()
        assert ref.scalars()["deltaT"] == orc.get("dt")
        assert_bitwise(ref.state(), orc.state(), "U after step %d" % (n + 1))
        assert_bitwise(ref.cell(10), orc.cell(10), "mu_SGS after step %d" % (n + 1))


@pytest.mark.parametrize("variant", VARIANTS)
def test_entry_points_bitwise_random_state(variant):
    ref, orc = pair(variant, 11, 10, 9, seed=4)
    ref.set_block(2, 7, 13.5, -4.0, 2.5)
    for k, v in (("k_min", 2), ("k_max", 7), ("Ud", 13.5), ("Vd", -4.0), ("Wd", 2.5)):
        orc.set(k, v)
    ref.ghost(); orc.ghost()
    assert_bitwise(ref.state(), orc.state(), "ghost fill")
    ref.recon(); orc.recon()
    for f in range(30):
        assert_bitwise(ref.face(f), orc.face(f), "recon face %d" % f)
    ref.iface(); orc.iface()
    for f in range(30):
        assert_bitwise(ref.face(f), orc.face(f), "bc face %d" % f)
    ref.fluxes(); orc.fluxes()
    for f in range(30):
        assert_bitwise(ref.face(f), orc.face(f), "flux face %d" % f)


@pytest.mark.parametrize("nstages", [2, 3])
def test_stage_count_variants(nstages):
    ref, orc = pair("dyn", 10, 10, 6, seed=6, nstages=nstages)
    for n in range(3):
        ref.step(); orc.step()
        assert_bitwise(ref.state(), orc.state(), "U after step %d" % (n + 1))
        assert ref.cur_is_p() == orc.cur_is_p()


def test_filter_and_random_and_minmod():
    need("cs")
    ref = RefLib("cs")
    ref.initialize(6, 5, 4)
    orc = Oracle(6, 5, 4)
    a = random_state(6, 5, 4, seed=9)[4]
    assert_bitwise(ref.filter(a), orc.filter(a), "filter_")
    x = X0
    for _ in range(100):
        assert ref.random(x) == Oracle.lib().or_random(x)
        x = ref.random(x)
    rng = np.random.default_rng(2)
    for x, y in rng.uniform(-3, 3, (500, 2)):
        assert ref.minmod(x, y) == Oracle.lib().or_minmod(x, y)

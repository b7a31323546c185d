"""The drop-in boundary driven the way the reference drives its hot path.

tests/layer2_dropin.c allocates the reference's real*** arrays with
Array3D, hands them to fsl_layer2_attach, and runs the loop of
src/layer2.c:141-185 through libfsl_layer2.so: checkCoNum, the block draw
with Random, fsl_layer2_set_block, then per stage BounCondInGhostCells /
Reconstruction / BounCondOnInterfaces / Fluxes / Evolution, and
fsl_layer2_sync_host before the (omitted) Probes.  Its host arrays after N
steps -- state, mu_SGS, deltaT, LCG and block -- must equal the oracle's bit
for bit.  Also filter_ (a new host array) and DynamicSmagorinsky's argument
check (the arrays must be the attached ones)."""
import os
import subprocess

import numpy as np
import pytest

from fsl_test_util import assert_bitwise, random_state
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_bin", "layer2_dropin")
SEED = 0.317391239817


def _run(args, timeout=300):
    r = subprocess.run([EXE] + [str(a) for a in args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=timeout)
    print(r.stdout[-400:], r.stderr[-400:])
    return r


def _read_run(path, L, H, D):
    G = (L + 2) * (H + 2) * (D + 2)
    raw = open(path, "rb").read()
    cells = np.frombuffer(raw[:6 * G * 8], dtype=np.float64).reshape(6, L + 2, H + 2, D + 2)
    tail = raw[6 * G * 8:]
    dt, X, Ud, Vd, Wd = np.frombuffer(tail[:40], dtype=np.float64)
    counter, kmin, kmax, st = np.frombuffer(tail[40:56], dtype=np.int32)
    return cells[:5], cells[5], dict(dt=dt, X=X, Ud=Ud, Vd=Vd, Wd=Wd, counter=counter, k_min=kmin,
                                      k_max=kmax, status=st)


@pytest.mark.parametrize("mesh,nsteps", [((16, 12, 8), 6), ((64, 48, 32), 3)])
@pytest.mark.parametrize("dynamic", [0, 1])
@pytest.mark.parametrize("start", ["reference", "random"])
def test_reference_call_order_matches_the_oracle(tmp_path, mesh, nsteps, dynamic, start):
    L, H, D = mesh
    orc = Oracle(L, H, D, dynamic=dynamic)
    orc.set("BL_HIG", min(5, H - H // 2))
    orc.set("nStages", 3)
    orc.set("X", SEED)
    if start == "reference":
        orc.ic_reference()
    else:
        orc.set_state(random_state(L, H, D, seed=7))
    U0 = orc.state()
    U0.tofile(str(tmp_path / "state.bin"))
    r = _run(["run", L, H, D, dynamic, 3, nsteps, "%.17g" % SEED, tmp_path / "state.bin", tmp_path / "out.bin"])
    assert r.returncode == 0, r.stderr
    U, mu, sc = _read_run(tmp_path / "out.bin", L, H, D)
    for _ in range(nsteps):
        orc.step()
    ref = orc.state()
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    assert_bitwise(ref[inner], U[inner], "state after %d steps" % nsteps)
    assert_bitwise(ref, U, "state with ghost layer")
    if dynamic:
        assert_bitwise(orc.cell(10), mu, "mu_SGS")
    assert sc["status"] == 0
    assert sc["dt"] == orc.get("dt")
    assert sc["X"] == orc.get("X")
    assert (sc["k_min"], sc["k_max"]) == (int(orc.get("k_min")), int(orc.get("k_max")))
    assert (sc["Ud"], sc["Vd"], sc["Wd"]) == (orc.get("Ud"), orc.get("Vd"), orc.get("Wd"))
    orc.close()


def test_two_stage_fall_through_through_the_drop_in(tmp_path):
    """nStages = 2 (layer2.ini:21 as shipped): Evolution runs the two-stage
    update and then the three-stage one (src/evolution.c:19-22)."""
    L, H, D = 16, 12, 8
    orc = Oracle(L, H, D, dynamic=1)
    orc.set("BL_HIG", 5)
    orc.set("nStages", 2)
    orc.set("X", SEED)
    orc.ic_reference()
    orc.state().tofile(str(tmp_path / "state.bin"))
    r = _run(["run", L, H, D, 1, 2, 4, "%.17g" % SEED, tmp_path / "state.bin", tmp_path / "out.bin"])
    assert r.returncode == 0, r.stderr
    U, _, sc = _read_run(tmp_path / "out.bin", L, H, D)
    for _ in range(4):
        orc.step()
    assert_bitwise(orc.state(), U, "2-stage state")
    assert sc["dt"] == orc.get("dt")
    orc.close()


def test_filter_returns_a_new_host_array(tmp_path):
    L, H, D = 16, 12, 8
    a = np.random.default_rng(2).uniform(-1, 1, (L + 2, H + 2, D + 2))
    a.tofile(str(tmp_path / "in.bin"))
    r = _run(["filter", L, H, D, tmp_path / "in.bin", tmp_path / "out.bin"])
    assert r.returncode == 0, r.stderr
    got = np.fromfile(str(tmp_path / "out.bin"), dtype=np.float64).reshape(a.shape)
    orc = Oracle(L, H, D, dynamic=1)
    assert_bitwise(orc.filter(a), got, "filter_")
    orc.close()


def test_dynamic_smagorinsky_needs_the_attached_arrays(tmp_path):
    L, H, D = 16, 12, 8
    U = random_state(L, H, D, seed=9)
    U.tofile(str(tmp_path / "state.bin"))
    r = _run(["dynarg", L, H, D, tmp_path / "state.bin", tmp_path / "out.bin"])
    assert r.returncode == 0, r.stderr
    raw = open(tmp_path / "out.bin", "rb").read()
    st = np.frombuffer(raw[:8], dtype=np.int32)
    mu = np.frombuffer(raw[8:], dtype=np.float64).reshape(L + 2, H + 2, D + 2)
    assert st[0] == 0                       # U1..U4, mu_SGS as attached
    assert st[1] == -6                      # FSL_ERR_UNSUPPORTED: other arrays
    orc = Oracle(L, H, D, dynamic=1)
    orc.set_state(U)
    orc.dynamic()
    assert_bitwise(orc.cell(10), mu, "mu_SGS of DynamicSmagorinsky()")
    orc.close()

"""The CPU oracle against the golden fixtures written from the compiled
reference (tests/golden/make_golden.py).  Bit-for-bit."""
import os

import numpy as np
import pytest

from fsl_test_util import assert_bitwise
from oracle_lib import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
X0 = 0.317391239817
FLUX_IDS = (0, 1, 2, 3, 4, 10, 11, 12, 13, 14, 20, 21, 22, 23, 24)


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def replay_stage1(orc, g):
    orc.check_co_num()
    orc.block_draw()
    assert orc.get("dt") == g["dt"]
    blk = [orc.get(k) for k in ("k_min", "k_max", "Ud", "Vd", "Wd", "X")]
    assert blk == list(g["block"])
    orc.ghost()
    assert_bitwise(g["U_ghost"], orc.state(), "state after ghost fill")
    orc.recon()
    assert_bitwise(g["faces_recon"], np.concatenate([orc.face(f).ravel() for f in range(30)]),
                   "faces after Reconstruction")
    orc.iface()
    assert_bitwise(g["faces_bc"], np.concatenate([orc.face(f).ravel() for f in range(30)]),
                   "faces after BounCondOnInterfaces")
    orc.fluxes()
    assert_bitwise(g["fluxes"], np.concatenate([orc.face(f).ravel() for f in FLUX_IDS]), "fluxes")
    assert_bitwise(g["mu_sgs"], orc.cell(10), "mu_SGS")
    orc.evolution(3, 1)
    assert_bitwise(g["Up_stage1"], np.stack([orc.cell(5 + f) for f in range(5)]), "Up after stage 1")
    for st in (2, 3):
        orc.ghost(); orc.recon(); orc.iface(); orc.fluxes(); orc.evolution(3, st)
        assert_bitwise(g["U_stage%d" % st], orc.state(), "U after stage %d" % st)
        assert_bitwise(g["Up_stage%d" % st], np.stack([orc.cell(5 + f) for f in range(5)]),
                       "Up after stage %d" % st)


@pytest.mark.parametrize("variant", ["cs", "dyn"])
def test_g1_reference_case_stage_by_stage(variant):
    g = load("g1_ref_" + variant)
    L, H, D = g["mesh"]
    orc = Oracle(L, H, D, dynamic=int(variant == "dyn"))
    orc.ic_reference()
    assert_bitwise(g["U0"], orc.state(), "initial condition")
    replay_stage1(orc, g)


@pytest.mark.parametrize("variant", ["cs", "dyn"])
def test_g2_random_state_stage_by_stage(variant):
    g = load("g2_rand_" + variant)
    L, H, D = g["mesh"]
    orc = Oracle(L, H, D, dynamic=int(variant == "dyn"))
    orc.set_state(g["U0"])
    replay_stage1(orc, g)
    for n in (2, 3):
        orc.step()
        assert orc.get("dt") == g["dt_step%d" % n]
        assert_bitwise(g["U_step%d" % n], orc.state(), "U after step %d" % n)


@pytest.mark.slow
@pytest.mark.parametrize("variant", ["cs", "dyn"])
def test_g3_hundred_steps(variant):
    g = load("g3_long_" + variant)
    L, H, D = g["mesh"]
    orc = Oracle(L, H, D, dynamic=int(variant == "dyn"))
    orc.ic_reference()
    for n in range(1, 101):
        orc.step()
        assert orc.get("dt") == g["dt"][n - 1], "dt at step %d" % n
        blk = [orc.get(k) for k in ("k_min", "k_max", "Ud", "Vd", "Wd", "counter")]
        assert blk == list(g["blocks"][n - 1]), "block at step %d" % n
        if n in (1, 10, 100):
            U = orc.state()
            Ui = U[:, 1:-1, 1:-1, 1:-1]
            norms = np.array([[np.abs(Ui[f]).sum(), np.sqrt((Ui[f] ** 2).sum()), np.abs(Ui[f]).max()]
                              for f in range(5)])
            assert_bitwise(g["norms_step%d" % n], norms, "norms after step %d" % n)
            k = D // 2 + 1
            assert_bitwise(g["xlines_step%d" % n], U[:, :, [H // 4, H // 2, 3 * H // 4], k],
                           "x-lines after step %d" % n)


def test_g4_lcg_and_blocks():
    g = load("g4_lcg")
    lib = Oracle.lib()
    seq = [X0]
    for _ in range(64):
        seq.append(lib.or_random(seq[-1]))
    assert_bitwise(g["random_seq"], np.array(seq), "Random sequence")
    for D in (32, 96, 256):
        orc = Oracle(4, 4, D)
        draws = []
        for _ in range(16):
            orc.set("counter", 0)
            orc.block_draw()
            draws.append([orc.get(k) for k in ("counter", "k_min", "k_max", "Ud", "Vd", "Wd", "X")])
        assert_bitwise(g["blocks_D%d" % D], np.array(draws), "block draws DEP=%d" % D)


def test_g5_filter():
    g = load("g5_filter")
    L, H, D = g["mesh"]
    orc = Oracle(L, H, D)
    assert_bitwise(g["output"], orc.filter(g["input"]), "filter_")


def test_g6_two_stage_fallthrough():
    g = load("g6_nstages2")
    L, H, D = g["mesh"]
    orc = Oracle(L, H, D, dynamic=1)
    orc.set("nStages", 2)
    orc.ic_reference()
    for n in range(1, 5):
        orc.step()
        assert orc.get("dt") == g["dt_step%d" % n]
        assert_bitwise(g["U_step%d" % n], orc.state(), "U after step %d" % n)
        assert_bitwise(g["Up_step%d" % n], np.stack([orc.cell(5 + f) for f in range(5)]),
                       "Up after step %d" % n)


def test_block_draw_golden():
    """First block of the reference run log for X0 and DEP = 32: counter 41,
    k_min 3, k_max 11, Ud 19.824909, Vd -1.837488, Wd -30.123856."""
    orc = Oracle(4, 4, 32)
    orc.block_draw()
    assert orc.get("counter") == 41 and orc.get("k_min") == 3 and orc.get("k_max") == 11
    assert round(orc.get("Ud"), 6) == 19.824909
    assert round(orc.get("Vd"), 6) == -1.837488
    assert round(orc.get("Wd"), 6) == -30.123856

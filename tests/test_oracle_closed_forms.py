"""Closed forms and invariants that pin the CPU oracle independently of the
reference's own numbers, chosen so that plausible mistakes (a dropped term,
a wrong sign, a transposed index) fail one of them."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from fsl_test_util import random_state
from oracle_lib import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def minmod_spec(x, y):
    """Chakravarthy-Osher minmod with compression 4: 0 for opposite signs or a
    zero argument, else the argument of x's sign with magnitude min(|x|, 4|y|)."""
    if x * y <= 0:
        return 0.0
    return math.copysign(min(abs(x), 4 * abs(y)), x)


@pytest.mark.parametrize("x,y", [(1.0, 2.0), (1.0, 0.1), (-3.0, -0.5), (-3.0, -2.0), (2.0, -1.0),
                                 (0.0, 5.0), (5.0, 0.0), (4.0, 1.0), (-4.0, -1.0), (1e-300, 1e-300),
                                 (7.5, 1.875), (-0.25, -1e-9)])
def test_minmod_truth_table(x, y):
    assert Oracle.lib().or_minmod(x, y) == minmod_spec(x, y)


def test_random_is_frac_of_11x_plus_pi():
    lib = Oracle.lib()
    x = 0.317391239817
    for _ in range(200):
        nx = lib.or_random(x)
        t = 11.0 * x + math.pi
        assert nx == t - float(int(t))
        assert 0.0 <= nx < 1.0
        x = nx


def uniform_oracle(dynamic, u=100.0, L=12, H=10, D=6):
    orc = Oracle(L, H, D, dynamic=dynamic)
    for k in ("Ua", "Va", "Wa"):
        orc.set(k, 0.0)
    orc.u_in()[:] = u
    U = np.zeros((5, L + 2, H + 2, D + 2))
    U[0] = 1.0
    U[1] = u
    U[4] = 1e5 / 0.4 + 0.5 * u * u
    orc.set_state(U)
    return orc, U


@pytest.mark.parametrize("dynamic", [0, 1])
def test_uniform_flow_is_a_fixed_point(dynamic):
    """A uniform stream matching the inflow, outflow pressure and slip walls:
    every face carries the same flux, so L(U) = 0 exactly (stage 1 returns U
    bit for bit); full steps drift only by the rounding of the RK weights
    1/3 + 2/3 != 1 (a few ulp)."""
    orc, U0 = uniform_oracle(dynamic)
    orc.check_co_num(); orc.block_draw(); orc.ghost(); orc.recon(); orc.iface(); orc.fluxes()
    orc.evolution(3, 1)
    Up = np.stack([orc.cell(5 + f) for f in range(5)])
    assert np.array_equal(Up[:, 1:-1, 1:-1, 1:-1], U0[:, 1:-1, 1:-1, 1:-1])
    orc2, _ = uniform_oracle(dynamic)
    for _ in range(3):
        orc2.step()
    U = orc2.state()[:, 1:-1, 1:-1, 1:-1]
    scale = [1.0, 100.0, 100.0, 100.0, np.abs(U0[4]).max()]
    for f in range(5):
        assert np.abs(U[f] - U0[f, 1:-1, 1:-1, 1:-1]).max() <= 1e-14 * scale[f]


def euler_flux(q, n):
    r = q[0]
    v = q[1:4] / r
    p = (q[4] - 0.5 * r * (v @ v)) * 0.4
    F = np.zeros(5)
    F[0] = r * v[n]
    F[1:4] = r * v[n] * v
    F[1 + n] += p
    F[4] = v[n] * (3.5 * p + 0.5 * r * (v @ v))
    return F


def prim_to_cons(r, u, v, w, p):
    return np.array([r, r * u, r * v, r * w, p / 0.4 + 0.5 * r * (u * u + v * v + w * w)])


def flux_of_face_states(left, right, n, dynamic=0):
    """Flux through every x/y/z face when the cells are uniform (viscous part
    zero) and the face states are given."""
    L, H, D = 6, 10, 5
    orc = Oracle(L, H, D, dynamic=dynamic)
    U = np.zeros((5, L + 2, H + 2, D + 2))
    U[0] = 1.0; U[1] = 10.0; U[4] = 1e5 / 0.4 + 50.0
    orc.set_state(U)
    base = 10 * n
    for f in range(5):
        orc.face(base + f)[...] = left[f]
        orc.face(base + 5 + f)[...] = right[f]
    orc.fluxes()
    return np.array([orc.face(base + f).ravel()[3] for f in range(5)])


@pytest.mark.parametrize("n", [0, 1, 2])
def test_riemann_supersonic_takes_left_state(n):
    c = math.sqrt(1.4 * 1e5 / 1.2)
    vel = [30.0, -20.0, 15.0]
    vel[n] = 2.0 * c
    left = prim_to_cons(1.2, *vel, 1e5)
    vr = list(vel); vr[n] = 2.5 * c
    right = prim_to_cons(0.9, *vr, 0.8e5)
    F = flux_of_face_states(left, right, n)
    np.testing.assert_allclose(F, euler_flux(left, n), rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("n", [0, 1, 2])
def test_riemann_mirror_states_stagnate(n):
    vel = [25.0, -35.0, 15.0]
    vel[n] = 40.0
    left = prim_to_cons(1.1, *vel, 1.05e5)
    vr = list(vel); vr[n] = -vel[n]
    right = prim_to_cons(1.1, *vr, 1.05e5)
    F = flux_of_face_states(left, right, n)
    # u* = 0: no mass or energy through the face; two colliding streams
    # compress, so the normal momentum flux p* exceeds p
    assert abs(F[0]) < 1e-9 * 40.0
    assert abs(F[4]) < 1e-9 * 40.0 * 1e5
    assert F[1 + n] > 1.05e5
    # expanding mirror pair: p* below p
    F2 = flux_of_face_states(right, left, n)
    assert abs(F2[0]) < 1e-9 * 40.0 and F2[1 + n] < 1.05e5


@pytest.mark.parametrize("n", [0, 1, 2])
def test_riemann_equal_states_give_euler_flux(n):
    q = prim_to_cons(0.95, 60.0, -25.0, 12.0, 0.97e5)
    F = flux_of_face_states(q, q, n)
    np.testing.assert_allclose(F, euler_flux(q, n), rtol=1e-12, atol=1e-8)


def test_rk3_reproduces_shu_osher_polynomial():
    """dU/dt = lambda U through Evolution(3, s): with x-fluxes chosen so that
    L(U_s) = z U_s at every stage, three stages give (1 + z + z^2/2 + z^3/6) U."""
    L, H, D = 8, 10, 4
    orc = Oracle(L, H, D, dynamic=0)
    U0 = random_state(L, H, D, seed=5)
    orc.set_state(U0)
    dt = 2e-4
    orc.set("dt", dt)
    for k, d in (("dtX", 1.6666667), ("dtY", 1.0), ("dtZ", 1.5)):
        orc.set(k, dt / d)
    dtX = orc.get("dtX")
    z = -0.37
    for fid in range(10, 30):
        orc.face(fid)[...] = 0.0
    stage_state = U0
    for s in (1, 2, 3):
        for f in range(5):
            Fx = orc.face(f)
            Fx[0] = 0.0
            for i in range(1, L + 1):
                # dtX (Fx[i-1] - Fx[i]) = z U_s(i)
                Fx[i] = Fx[i - 1] - z * stage_state[f, i, 1:-1, 1:-1] / dtX
        orc.evolution(3, s)
        stage_state = np.stack([orc.cell(5 + f) for f in range(5)]) if s < 3 else orc.state()
    poly = 1 + z + z * z / 2 + z ** 3 / 6
    Ui = orc.state()[:, 1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(Ui, poly * U0[:, 1:-1, 1:-1, 1:-1], rtol=1e-11)


def test_evolution_conserves_up_to_boundary_fluxes():
    """Flux form: the interior sum changes only by what crosses the domain
    faces (telescoping), for arbitrary fluxes."""
    L, H, D = 9, 10, 7
    orc = Oracle(L, H, D, dynamic=0)
    U0 = random_state(L, H, D, seed=8)
    orc.set_state(U0)
    rng = np.random.default_rng(9)
    for fid in (0, 1, 2, 3, 4, 10, 11, 12, 13, 14, 20, 21, 22, 23, 24):
        orc.face(fid)[...] = rng.uniform(-1e3, 1e3, orc.face(fid).shape)
    dtX, dtY, dtZ = orc.get("dtX"), orc.get("dtY"), orc.get("dtZ")
    expect = []
    for f in range(5):
        fx, fy, fz = orc.face(f), orc.face(10 + f), orc.face(20 + f)
        expect.append(dtX * (fx[0].sum() - fx[L].sum()) + dtY * (fy[:, 0].sum() - fy[:, H].sum())
                      + dtZ * (fz[:, :, 0].sum() - fz[:, :, D].sum()))
    orc.evolution(3, 1)
    Up = np.stack([orc.cell(5 + f) for f in range(5)])
    for f in range(5):
        got = Up[f, 1:-1, 1:-1, 1:-1].sum() - U0[f, 1:-1, 1:-1, 1:-1].sum()
        assert abs(got - expect[f]) < 1e-9 * max(1.0, np.abs(U0[f]).sum())


def test_filter_of_constant_field_and_its_truncations():
    """filter_ of a constant c: c in the interior; the x and z passes drop the
    missing neighbour at i = L and k = D (2c/3 each), j = 0 / H+1 read zero
    scratch, i = 0 keeps the y-pass value."""
    L, H, D = 7, 6, 5
    orc = Oracle(L, H, D)
    c = 3.0
    out = orc.filter(np.full((L + 2, H + 2, D + 2), c))
    np.testing.assert_allclose(out[1:L, 1:H + 1, 1:D], c, rtol=1e-15)
    np.testing.assert_allclose(out[1:L, 1:H + 1, D], 2 * c / 3, rtol=1e-15)
    np.testing.assert_allclose(out[L, 1:H + 1, 1:D], 2 * c / 3, rtol=1e-15)
    np.testing.assert_allclose(out[L, 1:H + 1, D], 4 * c / 9, rtol=1e-15)
    np.testing.assert_allclose(out[0, 1:H + 1, 0:D + 1], c, rtol=1e-15)
    assert np.all(out[:, 0, :][1:] == 0.0)
    assert np.all(out[:, H + 1, :] == 0.0)
    assert np.all(out[L + 1] == 0.0)
    assert np.all(out[:, :, D + 1] == 0.0)
    assert np.all(out[1:, :, 0] == 0.0)


def test_dynamic_model_vanishes_for_uniform_translation():
    orc, _ = uniform_oracle(1)
    orc.ghost()
    orc.dynamic()
    assert np.all(orc.cell(10) == 0.0)


def test_planted_fault_is_caught_by_golden():
    """Dropping the heat-flux term in the oracle must break the golden flux
    comparison (the parity checks are sensitive to a dropped term)."""
    g = np.load(os.path.join(GOLD, "g2_rand_cs.npz"), allow_pickle=False)
    L, H, D = g["mesh"]
    flag = C.c_int.in_dll(Oracle.lib(), "or_fault_drop_heat_flux")
    flag.value = 1
    try:
        orc = Oracle(L, H, D, dynamic=0)
        orc.set_state(g["U0"])
        orc.check_co_num(); orc.block_draw(); orc.ghost(); orc.recon(); orc.iface(); orc.fluxes()
        got = np.concatenate([orc.face(f).ravel() for f in (0, 1, 2, 3, 4, 10, 11, 12, 13, 14,
                                                             20, 21, 22, 23, 24)])
        assert not np.array_equal(got, g["fluxes"])
    finally:
        flag.value = 0


def test_z_quirk_is_what_the_reference_does():
    """With the z quirk off the reconstructed z faces differ from the golden
    (the v row of the z faces uses W*ee in the reference)."""
    g = np.load(os.path.join(GOLD, "g2_rand_dyn.npz"), allow_pickle=False)
    L, H, D = g["mesh"]
    orc = Oracle(L, H, D, dynamic=1)
    orc.set("z_quirk", 0)
    orc.set_state(g["U0"])
    orc.check_co_num(); orc.block_draw(); orc.ghost(); orc.recon()
    got = np.concatenate([orc.face(f).ravel() for f in range(30)])
    assert not np.array_equal(got, g["faces_recon"])


def test_dynamic_terms_reproduce_the_pointwise_model():
    """The oracle's Germano terms (the test reference of the averaged
    coefficients) give its pointwise mu_SGS: clip(-L:M / 2 (M:M + 1e-30), 0,
    0.15) rho DD |S| (to rounding: the pointwise model associates rho Cd DD |S|)."""
    L, H, D = 16, 12, 8
    o = Oracle(L, H, D, dynamic=1)
    o.set_state(random_state(L, H, D, seed=41))
    o.ghost()
    o.dynamic()
    mu = o.cell(10).copy()
    lm, mm, fac = o.dynamic_terms()
    Cd = np.clip(-0.5 * (lm / (mm + 1e-30)), 0.0, 0.15)
    I = (slice(1, -1),) * 3
    assert np.abs(mu[I] - Cd[I] * fac[I]).max() <= 1e-14 * np.abs(mu).max()
    assert (Cd[I] > 0).any() and (Cd[I] < 0.15).any() and (Cd[I] == 0).any()
    assert np.all(lm[0] == 0) and np.all(fac[:, 0] == 0)    # ghost layer untouched

"""The GPU configurations of the benchmark case at their full sizes, through
size-independent properties (the oracle is far too slow at these sizes; the
FP64 path is bit-identical to it on every mesh the other tests check).

* 384x256x256, mixed precision against the FP64 dynamic path, Mc 0.3 and 0.7,
  20 steps: per-field max-norm differences, the time step, the mass and
  energy totals, momentum thickness delta_theta and integrated resolved TKE.
* 384x256x256, the learned closure, Mc 0.4: after the spin-up (the network
  leaves the dynamic fallback once a judge passes) it runs on the network,
  the next judge passes (error < 0.10) and the state stays finite.
* 512x384x384, the learned closure, Mc 0.6, Re 2000, 3 % perturbations:
  network steps after the spin-up, finite state and conserved totals.

Tolerances of the mixed path: about ten times the differences measured
after 20 steps on MI355X (noted beside each bound)."""
import os

import numpy as np
import pytest

import freeshearlayer_amd as fsl
from freeshearlayer_amd import shear_case

pytestmark = pytest.mark.gpu

MODEL = os.path.join(os.path.dirname(fsl.LIB_DIR), "data", "sgs_mlp.bin")
FULL = (384, 256, 256)


def rel_err(a, b):
    scale = [np.abs(b[0]).max(), np.abs(b[1]).max(), np.abs(b[1]).max(), np.abs(b[1]).max(), np.abs(b[4]).max()]
    return [float(np.abs(a[f] - b[f]).max() / scale[f]) for f in range(5)]


def leg(precision, mach, steps):
    ctx, case, U0, u_in, _ = shear_case.make(*FULL, sgs="dynamic", precision=precision, fused=True, cfl=0.5,
                                             mach=mach)
    try:
        ctx.run(steps)
        out = dict(U=ctx.download_state()[:, 1:-1, 1:-1, 1:-1], sums=np.array(ctx.conservation_sums()), dt=ctx.dt,
                   stats=shear_case.statistics(ctx, case, u_in))
    finally:
        ctx.close()
    return out


@pytest.mark.parametrize("mach", [0.3, 0.7])
def test_mixed_tracks_fp64_at_384x256x256(mach):
    a = leg("mixed", mach, 20)
    b = leg("fp64", mach, 20)
    e = rel_err(a["U"], b["U"])
    ds = np.abs(a["sums"] - b["sums"]) / np.abs(b["sums"])
    th = abs(a["stats"]["delta_theta"] - b["stats"]["delta_theta"]) / b["stats"]["delta_theta"]
    tk = abs(a["stats"]["tke"] - b["stats"]["tke"]) / b["stats"]["tke"]
    print("Mc %.1f, 20 steps: fields %s, sums %s, dtheta %.2e, tke %.2e, dt %.3e"
          % (mach, ["%.1e" % x for x in e], ["%.1e" % x for x in ds], th, tk, abs(a["dt"] - b["dt"]) / b["dt"]))
    assert np.all(np.isfinite(a["U"]))
    assert max(e) < 2e-5                            # measured 3e-7 (Mc 0.3), 2e-6 (Mc 0.7)
    assert ds[0] < 1e-9 and ds[4] < 1e-9            # mass and energy totals: 3e-11, 2e-10
    assert th < 1e-8 and tk < 1e-6                  # 2e-10 and 7e-8
    assert abs(a["dt"] - b["dt"]) / b["dt"] < 1e-6  # 5e-8


def spun_up_learned(mesh, mach, reynolds=500.0, pert=0.02, cap=300):
    ctx, case, U0, u_in, _ = shear_case.make(*mesh, sgs="learned", precision="mixed", fused=True, cfl=0.5,
                                             mach=mach, reynolds=reynolds, pert=pert)
    ctx.mlp_load(MODEL)
    ctx.learned_configure(50, 0.10)
    steps = 0
    while True:
        st = ctx.learned_status()
        if st["judged"] >= 1 and not st["in_fallback"] and st["learned_steps"] > 0:
            break
        assert steps < cap, "still on the dynamic fallback after %d steps: %s" % (steps, st)
        ctx.step()
        steps += 1
    return ctx, case, u_in, steps


def test_learned_closure_runs_on_the_network_at_384x256x256():
    ctx, case, u_in, spun = spun_up_learned(FULL, 0.4)
    try:
        before = ctx.learned_status()
        ctx.run(60)                                   # contains the next judge
        st = ctx.learned_status()
        print("spin-up %d steps; status %s" % (spun, st))
        assert st["judged"] == before["judged"] + 1
        assert st["fallbacks"] == before["fallbacks"]
        assert st["learned_steps"] - before["learned_steps"] == 60
        assert 0 <= st["last_error"] < 0.10
        assert np.all(np.isfinite(ctx.conservation_sums()))
        s = shear_case.statistics(ctx, case, u_in)
        assert s["delta_theta"] > 0 and s["tke"] > 0
    finally:
        ctx.close()


def test_learned_closure_at_512x384x384():
    ctx, case, u_in, spun = spun_up_learned((512, 384, 384), 0.6, reynolds=2000.0, pert=0.03)
    try:
        before = ctx.learned_status()
        s0 = np.array(ctx.conservation_sums())
        ctx.run(5)
        st = ctx.learned_status()
        s1 = np.array(ctx.conservation_sums())
        print("spin-up %d steps; status %s" % (spun, st))
        assert st["learned_steps"] - before["learned_steps"] == 5
        assert np.all(np.isfinite(s1))
        # 5 steps move the totals only through the inflow / outflow faces
        assert np.all(np.abs(s1[[0, 4]] - s0[[0, 4]]) / np.abs(s0[[0, 4]]) < 1e-3)
    finally:
        ctx.close()

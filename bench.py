"""Throughput benchmark of the per-step update on one MI355X.

Workload (default): the GPU FP64 plain method with the dynamic Smagorinsky
model on the synthetic free-shear-layer case at 192x128x96 cells -- tanh
mean profile U(y) = Um + dU/2 tanh(2y/delta_w0), velocity ratio U2/U1 = 0.5,
convective Mach 0.3, Re 500 on delta_w0 = 4 dy, seeded 2 % dU perturbations
where |y| < delta_w0, acoustic CFL 0.5, 3-stage TVD RK.  Inputs are resident
in HBM before the timed region.

Metric: Mcell-updates/s per RK stage = cells x stages x steps / seconds.

Prints one JSON line (see the keys below).  `roofline` prices the dominant
kernel's algorithmic bytes per launch against its average launch duration,
both measured with HIP events on the context's stream over the timed steps.
`cpu_baseline` times the FP64 CPU oracle on one host core on a bounded
sample (two steps of the same workload) -- a reported baseline, not the
target.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20] [--mesh 192x128x96]
                    [--sgs dynamic|constant] [--precision fp64|mixed]
                    [--path fused|unfused] [--no-cpu]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import freeshearlayer_amd as fsl  # noqa: E402

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kernel, L, H, D, sgs):
    """Compulsory HBM bytes of one launch of `kernel` (FP64 state):
    each cell-sized array streamed once per access pattern it needs."""
    N = L * H * D
    G = (L + 2) * (H + 2) * (D + 2)
    nx, ny, nz = (L + 1) * H * D, L * (H + 1) * D, L * H * (D + 1)
    per = {
        # read U (5 fields), write 30 face states
        "recon": 5 * 8 * N + 30 * 8 * (nx + ny + nz) / 3,
        # read 10 face states + U (5 fields) + mu, write 5 fluxes
        "flux_x": 10 * 8 * nx + 5 * 8 * G + (8 * G if sgs else 0) + 5 * 8 * nx,
        "flux_y": 10 * 8 * ny + 5 * 8 * G + (8 * G if sgs else 0) + 5 * 8 * ny,
        "flux_z": 10 * 8 * nz + 5 * 8 * G + (8 * G if sgs else 0) + 5 * 8 * nz,
        # read 15 flux arrays, U and Up, write one state
        "rk_update": 5 * 8 * (nx + ny + nz) + 3 * 5 * 8 * N,
        # fused stage, averaged over the three stages: read the stage state
        # (40 B/cell), the step state in stages 2-3 (40 B/cell), mu_SGS (8) and
        # write the new state (40): (88 + 128 + 128) / 3 dynamic, (80+120+120)/3 constant
        "fused_stage": N * ((40 + 40 + (8 if sgs else 0)) + 2 * (40 + 40 + 40 + (8 if sgs else 0))) / 3.0,
        "filter_x": 2 * 8 * G, "filter_y": 2 * 8 * G, "filter_z": 2 * 8 * G,
    }
    return per.get(kernel)


def make_context(L, H, D, sgs, precision, path, cfl):
    case = fsl.case_reference(L, H, D)
    shear = fsl.shear_defaults()
    U, u_in, mu_L = fsl.ic_shear_layer(case, shear)
    case.mu_L = mu_L
    case.Pr_L = 0.72
    case.nStages = 3
    case.BL_HIG = min(5, H - H // 2)
    cfg = fsl.config(sgs_model=fsl.FSL_SGS_DYNAMIC if sgs == "dynamic" else fsl.FSL_SGS_CONSTANT,
                     precision=fsl.FSL_PREC_FP64 if precision == "fp64" else fsl.FSL_PREC_MIXED,
                     cfl_mode=fsl.FSL_CFL_ACOUSTIC, cfl=cfl, fused=1 if path == "fused" else 0,
                     rho_in=shear.rho0, p_in=shear.p0, p_out=shear.p0)
    ctx = fsl.Context(case, cfg)
    ctx.set_inflow(u_in)
    ctx.upload_state(U)
    ctx.set_lcg(0.317391239817, 0)
    return ctx, case, U, u_in


def cpu_baseline(L, H, D, sgs, U, u_in, case, steps=2):
    """FP64 CPU oracle (oracle/liboracle.so, scalar C, one core) on two steps
    of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    orc = Oracle(L, H, D, dynamic=int(sgs == "dynamic"))
    orc.set("mu_L", case.mu_L)
    orc.set("BL_HIG", case.BL_HIG)
    orc.u_in()[:] = u_in
    orc.derive()
    orc.set_state(U)
    orc.step()                       # first touch of every array
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.step()
    t = time.perf_counter() - t0
    orc.close()
    return {"value": L * H * D * 3 * steps / t / 1e6, "unit": "Mcell-updates/s per RK stage",
            "cores": 1, "kind": "port",
            "sample": "%d steps of the FP64 %s-SGS CPU oracle at %dx%dx%d (%.1f s)"
                      % (steps, sgs, L, H, D, t),
            "cpu": platform.processor() or platform.machine()}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mesh", default="192x128x96")
    ap.add_argument("--sgs", default="dynamic", choices=["dynamic", "constant"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--path", default="fused", choices=["fused", "unfused"])
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="N",
                    help="only run a plain device copy of N doubles 5 times (HBM counter calibration)")
    a = ap.parse_args(argv)
    if a.copy_calibration:
        ctx = fsl.Context(fsl.case_reference(8, 10, 8), fsl.config(sgs_model=fsl.FSL_SGS_CONSTANT))
        ctx.profile(True)
        ctx.copy_kernel(a.copy_calibration, 5)
        prof = ctx.profile_read()
        ms, cnt = prof["copy"]
        n = a.copy_calibration
        print(json.dumps({"copy_doubles": n, "bytes_read": 8 * n, "bytes_written": 8 * n,
                          "avg_us": ms / cnt * 1e3, "GBps": 16 * n / (ms / cnt / 1e3) / 1e9}))
        ctx.close()
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; the domain is not sharded across GPUs "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        return 2
    L, H, D = (int(v) for v in a.mesh.lower().split("x"))
    ctx, case, U0, u_in = make_context(L, H, D, a.sgs, a.precision, a.path, a.cfl)
    for _ in range(a.warmup):
        ctx.step()
    ctx.synchronize()
    ctx.profile_reset()
    if not a.no_profile:
        ctx.profile(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(a.steps)
    ctx.synchronize()
    t = time.perf_counter() - t0
    ctx.profile(False)
    prof = ctx.profile_read()
    sums = ctx.conservation_sums()
    if not np.all(np.isfinite(sums)):
        print("bench.py: non-finite state after the timed steps", file=sys.stderr)
        return 1
    N = L * H * D
    value = N * 3 * a.steps / t / 1e6
    roofline = None
    if prof:
        dom = max(prof.items(), key=lambda kv: kv[1][0])
        name, (ms, cnt) = dom
        avg_s = ms / cnt / 1e3
        b = algorithmic_bytes(name, L, H, D, a.sgs == "dynamic")
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
        wl = "%s %s-SGS %s path, %dx%dx%d cells, 3-stage TVD RK" % (a.precision.upper(), a.sgs, a.path,
                                                                     L, H, D)
        if os.path.exists(tpath):
            traffic = json.load(open(tpath)).get("%s@%s" % (name, wl))
        if b is not None:
            ach = b / avg_s / 1e9
            roofline = {"bound": "hbm", "kernel": name, "achieved": round(ach, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": traffic, "avg_launch_us": round(avg_s * 1e6, 2),
                        "alg_bytes_per_launch": int(b)}
        kernels = {k: {"ms_total": round(v[0], 3), "launches": int(v[1]),
                       "avg_us": round(v[0] / v[1] * 1e3, 2)} for k, v in sorted(prof.items())}
    else:
        kernels = {}
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(L, H, D, a.sgs, U0, u_in, case)
    line = {
        "metric": "Mcell-updates/s per RK stage on 1 MI355X, FP32+learned SGS vs FP64 dyn-Smag",
        "value": round(value, 2),
        "unit": "Mcell-updates/s per RK stage",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(t / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic shear layer (tanh profile, Mc 0.3, U2/U1 0.5, Re 500, 2%% seeded "
                "perturbations, acoustic CFL %.2f)" % a.cfl,
        "config": {"workload": "%s %s-SGS %s path, %dx%dx%d cells, 3-stage TVD RK"
                               % (a.precision.upper(), a.sgs, a.path, L, H, D),
                   "mesh": [L, H, D], "stages": 3, "sgs": a.sgs, "precision": a.precision,
                   "path": a.path},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "final_dt": ctx.dt,
        "mass_sum": sums[0],
    }
    print(json.dumps(line))
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Throughput benchmark of the per-step update on one MI355X.

Default workload: the FP32 + learned-closure path (mixed precision: FP32
stencil arithmetic, FP64 state / pressure recovery / RK / CFL; the per-cell
network for Cs^2 on the matrix cores (FP16 operands, FP32 accumulation;
trained in BF16), judged against the dynamic model every 50
steps with a dynamic fallback above 10 % error) on the synthetic free shear
layer at 384x256x256 cells, Mc 0.4 -- tanh mean profile U(y) = Um + dU/2
tanh(2y/delta_w0), U2/U1 = 0.5, Re 500 on delta_w0 = 4 dy, seeded 2 % dU
perturbations where |y| < delta_w0, acoustic CFL 0.5, 3-stage TVD RK --
timed beside the plain FP64 dynamic-Smagorinsky path on the same grid and
initial state.  Inputs are resident in HBM before each timed region; the
120 untimed warm-up steps of each leg carry the run past the judges of the
synthetic initial state (laminar profile plus white noise), which the
learned closure leaves to its dynamic fallback.

Metric: Mcell-updates/s per RK stage = cells x stages x steps / seconds.

Prints one JSON line.  `value` is the headline path's throughput;
`fp64_dynamic` the comparison leg and `speedup_vs_fp64_dynamic` their
ratio.  `roofline` prices the dominant kernel's algorithmic bytes per
launch against its average launch duration, measured with HIP events on the
context's stream over the timed steps; `mlp_roofline` the learned-closure
kernel's algorithmic FLOPs against the FP16 dense peak (its products are FP16).  `traffic` (HBM bytes per
launch) and the VALU / MFMA busy fractions come from rocprofv3 counter runs
of the same command (profiles/).  `cpu_baseline` times the reference's
own serial path (src/, compiled from its sources into
oracle/_ref/libfslref_dyn.so with real = double and zero-initialised
arrays) on one host core and a bounded sample, or, where that library was
not built, the FP64 CPU oracle (the same algorithm restated in C; also
reported as `cpu_port`) -- a reported baseline, not the target.

    python bench.py [--gpus 1] [--steps 300] [--warmup 120] [--mesh 384x256x256]
                    [--sgs learned|dynamic|constant] [--precision mixed|fp64]
                    [--mach 0.4] [--path fused|unfused] [--no-compare] [--no-cpu]
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
from freeshearlayer_amd import shear_case  # noqa: E402

HBM_PEAK_GBS = 8000.0
F16_PEAK_TFLOPS = 2500.0   # MI355X dense FP16 / BF16 MFMA
MODEL = os.path.join(ROOT, "freeshearlayer_amd", "data", "sgs_mlp.bin")
METRIC = "Mcell-updates/s per RK stage on 1 MI355X, FP32+learned SGS vs FP64 dyn-Smag"
UNIT = "Mcell-updates/s per RK stage"


def algorithmic_bytes(kernel, L, H, D, sgs):
    """Compulsory HBM bytes of one launch of `kernel` (FP64 state): each
    cell-sized array streamed once per access pattern it needs."""
    N = L * H * D
    G = (L + 2) * (H + 2) * (D + 2)
    nx, ny, nz = (L + 1) * H * D, L * (H + 1) * D, L * H * (D + 1)
    mu = 8 if sgs else 0
    per = {
        # read U (5 fields), write 30 face states
        "recon": 5 * 8 * N + 30 * 8 * (nx + ny + nz) / 3,
        # read 10 face states + U (5 fields) + mu, write 5 fluxes
        "flux_x": 10 * 8 * nx + 5 * 8 * G + (8 * G if sgs else 0) + 5 * 8 * nx,
        "flux_y": 10 * 8 * ny + 5 * 8 * G + (8 * G if sgs else 0) + 5 * 8 * ny,
        "flux_z": 10 * 8 * nz + 5 * 8 * G + (8 * G if sgs else 0) + 5 * 8 * nz,
        # read 15 flux arrays, U and Up, write one state
        "rk_update": 5 * 8 * (nx + ny + nz) + 3 * 5 * 8 * N,
        # fused stage (FP64 or mixed: both stream FP64 state), averaged over
        # the three stages: read the stage state (40 B/cell), the step state
        # in stages 2-3 (40), mu_SGS (8), write the new state (40)
        "fused_stage": N * ((40 + 40 + mu) + 2 * (40 + 40 + 40 + mu)) / 3.0,
        "filter_x": 2 * 8 * G, "filter_y": 2 * 8 * G, "filter_z": 2 * 8 * G,
    }
    return per.get(kernel)


# Learned closure per cell: 4 672 multiply-adds of the network (8x64 + 64x64
# + 64x1) = 9 344 algorithmic FLOPs; the MFMA tiles execute 12 288 (K of the
# first layer padded 8 -> 16, w3 in column 0 of a 16-wide tile)
MLP_FLOP_PER_CELL = 2 * (8 * 64 + 64 * 64 + 64 * 1)
MLP_MFMA_FLOP_PER_CELL = 2 * (16 * 64 + 64 * 64 + 64 * 16)


def host_cpu():
    """CPU model and logical core count of the host."""
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def workload_name(precision, sgs, path, L, H, D, mach):
    return "%s %s-SGS %s path, %dx%dx%d cells, Mc %.1f, 3-stage TVD RK" % (
        "FP32+FP64" if precision == "mixed" else "FP64", sgs, path, L, H, D, mach)


def time_leg(L, H, D, sgs, precision, path, cfl, mach, steps, warmup, profile=True):
    ctx, case, U0, u_in, _ = shear_case.make(L, H, D, sgs=sgs, precision=precision, fused=(path == "fused"),
                                             cfl=cfl, mach=mach)
    if sgs == "learned":
        p, n = fsl.mlp_file_read(MODEL)
        ctx.mlp_set_params(p, n)
        ctx.learned_configure(50, 0.10)
    for _ in range(warmup):
        ctx.step()
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profile(profile)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(steps)
    ctx.synchronize()
    t = time.perf_counter() - t0
    ctx.profile(False)
    prof = ctx.profile_read()
    sums = ctx.conservation_sums()
    status = ctx.learned_status() if sgs == "learned" else None
    out = {"ctx": ctx, "case": case, "U0": U0, "u_in": u_in, "t": t, "prof": prof, "sums": sums,
           "status": status, "value": L * H * D * 3 * steps / t / 1e6}
    return out


def roofline_of(prof, L, H, D, sgs, workload):
    if not prof:
        return None, {}
    kernels = {k: {"ms_total": round(v[0], 3), "launches": int(v[1]), "avg_us": round(v[0] / v[1] * 1e3, 2)}
               for k, v in sorted(prof.items())}
    name, (ms, cnt) = max(prof.items(), key=lambda kv: kv[1][0])
    b = algorithmic_bytes(name, L, H, D, sgs != "constant")
    if b is None:
        return None, kernels
    avg_s = ms / cnt / 1e3
    prof_tab = profile_table()
    traffic = prof_tab.get("%s@%s" % (name, workload))
    ach = b / avg_s / 1e9
    return ({"bound": "hbm", "kernel": name, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "avg_launch_us": round(avg_s * 1e6, 2),
             "alg_bytes_per_launch": int(b),
             "valu_busy_pct": prof_tab.get("valu_busy@%s@%s" % (name, workload))}, kernels)


def profile_table():
    """Per-kernel figures measured with rocprofv3 counters (profiles/)."""
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    return json.load(open(tpath)) if os.path.exists(tpath) else {}


def cpu_baseline(steps=2, mesh=(192, 128, 96)):
    """The FP64 dynamic-SGS CPU oracle (oracle/liboracle.so: the reference's
    serial algorithm in scalar C, one core) on a bounded sample: `steps`
    steps of the same shear layer on a smaller mesh (the per-cell cost does
    not depend on the mesh size)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    L, H, D = mesh
    c, U, u_in, _ = shear_case.initial(L, H, D, mach=0.4)
    orc = Oracle(L, H, D, dynamic=1)
    orc.set("mu_L", c.mu_L)
    orc.set("BL_HIG", c.BL_HIG)
    orc.u_in()[:] = u_in
    orc.derive()
    orc.set_state(U)
    orc.step()                       # first touch of every array
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.step()
    t = time.perf_counter() - t0
    orc.close()
    model, ncpu = host_cpu()
    return {"value": round(L * H * D * 3 * steps / t / 1e6, 4), "unit": UNIT, "cores": 1, "kind": "port",
            "sample": "%d steps of the FP64 dynamic-SGS CPU oracle on the same shear layer at %dx%dx%d "
                      "(%.1f s)" % (steps, L, H, D, t),
            "cpu": model, "host_logical_cores": ncpu}


def cpu_reference(steps=4, mesh=(192, 128, 96)):
    """The compiled reference (oracle/_ref/libfslref_dyn.so: src/ with real =
    double, FP64 dynamic Smagorinsky, rows zero-initialised) stepping the same
    shear-layer state on one core: `steps` of its `layer2` steps (checkCoNum,
    block draw, 3 stages of ghost fill, reconstruction, interface states,
    fluxes with the dynamic model, evolution).  Its own inflow and time-step
    rule apply; the per-cell cost does not depend on them.  None when the
    library was not built."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import RefLib, ref_available
    if not ref_available("dyn"):
        return None
    L, H, D = mesh
    c, U, _, _ = shear_case.initial(L, H, D, mach=0.4)
    ref = RefLib("dyn")
    ref.initialize(L, H, D, mu_L="%.17g" % c.mu_L, BL_HIG=c.BL_HIG)
    ref.set_state(U)
    ref.step()                       # first touch of every array
    t0 = time.perf_counter()
    for _ in range(steps):
        ref.step()
    t = time.perf_counter() - t0
    model, ncpu = host_cpu()
    return {"value": round(L * H * D * 3 * steps / t / 1e6, 4), "unit": UNIT, "cores": 1, "kind": "reference",
            "sample": "%d steps of the reference's serial src/ path (compiled from its sources, FP64, dynamic "
                      "SGS) on the same shear layer at %dx%dx%d (%.1f s)" % (steps, L, H, D, t),
            "cpu": model, "host_logical_cores": ncpu}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    # untimed: past the judges of the synthetic initial state (laminar profile
    # + white noise), which the learned closure leaves to the dynamic fallback
    ap.add_argument("--warmup", type=int, default=120)
    ap.add_argument("--mesh", default="384x256x256")
    ap.add_argument("--sgs", default="learned", choices=["learned", "dynamic", "constant"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--path", default="fused", choices=["fused", "unfused"])
    ap.add_argument("--mach", type=float, default=0.4)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 dynamic comparison leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="N",
                    help="only run a plain device copy of N doubles 5 times (HBM counter calibration)")
    a = ap.parse_args(argv)
    if a.copy_calibration:
        ctx = fsl.Context(fsl.case_reference(8, 10, 8), fsl.config(sgs_model=fsl.FSL_SGS_CONSTANT))
        ctx.profile(True)
        ctx.copy_kernel(a.copy_calibration, 5)
        ms, cnt = ctx.profile_read()["copy"]
        n = a.copy_calibration
        print(json.dumps({"copy_doubles": n, "bytes_read": 8 * n, "bytes_written": 8 * n,
                          "avg_us": ms / cnt * 1e3, "GBps": 16 * n / (ms / cnt / 1e3) / 1e9}))
        ctx.close()
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the per-step update runs on one GPU; the domain is not sharded across GPUs "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        return 2
    L, H, D = shear_case.parse_mesh(a.mesh)
    wl = workload_name(a.precision, a.sgs, a.path, L, H, D, a.mach)
    print("[bench] %s: %d warmup + %d timed steps" % (wl, a.warmup, a.steps), file=sys.stderr, flush=True)
    leg = time_leg(L, H, D, a.sgs, a.precision, a.path, a.cfl, a.mach, a.steps, a.warmup, not a.no_profile)
    ctx = leg["ctx"]
    if not np.all(np.isfinite(leg["sums"])):
        print("bench.py: non-finite state after the timed steps", file=sys.stderr)
        return 1
    roofline, kernels = roofline_of(leg["prof"], L, H, D, a.sgs, wl)
    mlp_roofline = None
    if "mlp_mu" in leg["prof"]:
        ms, cnt = leg["prof"]["mlp_mu"]
        tf = L * H * D * MLP_FLOP_PER_CELL / (ms / cnt / 1e3) / 1e12
        tab = profile_table()
        mlp_roofline = {"bound": "mfma", "kernel": "mlp_mu", "achieved": round(tf, 2), "peak": F16_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(tf / F16_PEAK_TFLOPS, 5),
                        "flop_per_cell": MLP_FLOP_PER_CELL, "mfma_flop_per_cell": MLP_MFMA_FLOP_PER_CELL,
                        "avg_launch_us": round(ms / cnt * 1e3, 2),
                        "mfma_busy_pct": tab.get("mfma_busy@mlp_mu@%s" % wl),
                        "valu_busy_pct": tab.get("valu_busy@mlp_mu@%s" % wl)}
    final_dt = ctx.dt
    ctx.close()
    compare = None
    if not a.no_compare and not (a.sgs == "dynamic" and a.precision == "fp64"):
        print("[bench] comparison leg: FP64 dynamic-SGS, same grid", file=sys.stderr, flush=True)
        ref = time_leg(L, H, D, "dynamic", "fp64", a.path, a.cfl, a.mach, a.steps, a.warmup, not a.no_profile)
        rl, _ = roofline_of(ref["prof"], L, H, D, "dynamic",
                            workload_name("fp64", "dynamic", a.path, L, H, D, a.mach))
        compare = {"value": round(ref["value"], 2), "ms_per_step": round(ref["t"] / a.steps * 1e3, 4),
                   "workload": workload_name("fp64", "dynamic", a.path, L, H, D, a.mach), "roofline": rl}
        ref["ctx"].close()
    cpu = port = None
    if not a.no_cpu:
        print("[bench] CPU reference sample", file=sys.stderr, flush=True)
        cpu = cpu_reference()
        print("[bench] CPU oracle sample", file=sys.stderr, flush=True)
        port = cpu_baseline()
        if cpu is None:
            cpu = port
    line = {
        "metric": METRIC,
        "value": round(leg["value"], 2),
        "unit": UNIT,
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(leg["t"] / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64+f16" if a.sgs == "learned" else "f32+f64",
        "data": "synthetic shear layer (tanh profile, Mc %.1f, U2/U1 0.5, Re 500, 2%% seeded perturbations, "
                "acoustic CFL %.2f)" % (a.mach, a.cfl),
        "config": {"workload": wl, "mesh": [L, H, D], "stages": 3, "sgs": a.sgs, "precision": a.precision,
                   "path": a.path, "mach": a.mach},
        "roofline": roofline,
        "mlp_roofline": mlp_roofline,
        "fp64_dynamic": compare,
        "speedup_vs_fp64_dynamic": round(leg["value"] / compare["value"], 3) if compare else None,
        "learned_status": leg["status"],
        "cpu_baseline": cpu,
        "cpu_port": port,
        "kernels": kernels,
        "final_dt": final_dt,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())

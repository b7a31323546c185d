"""Throughput benchmark of the per-step update on one MI355X.

Default workload: the FP32 + learned-closure path (mixed precision: FP32
stencil arithmetic, FP64 state / pressure recovery / RK / CFL; the per-cell
network for Cs^2 -- the shipped model, ReLU hidden layers -- on the matrix
cores (BF16 operands, FP32 accumulation, as in training), judged against the dynamic model every 50
steps with a dynamic fallback above 10 % error) on the synthetic free shear
layer at 384x256x256 cells, Mc 0.4 -- tanh mean profile U(y) = Um + dU/2
tanh(2y/delta_w0), U2/U1 = 0.5, Re 500 on delta_w0 = 4 dy, seeded 2 % dU
perturbations where |y| < delta_w0, acoustic CFL 0.5, 3-stage TVD RK --
timed beside the plain FP64 dynamic-Smagorinsky path on the same grid and
initial state.  Inputs are resident in HBM before each timed region.

Before its --warmup steps the learned leg runs an untimed spin-up that does
not depend on --warmup or --steps: it steps until a judge has passed and the
closure runs on the network (at most --spinup-cap steps).  Its learned-closure
counters are read before and after the timed steps; the timed window's own
counts are reported (`learned_window`), and if any timed step ran on the
dynamic fallback the line carries value null and the exit status is 1: the
headline is only ever a measurement of the network path.

Metric: Mcell-updates/s per RK stage = cells x stages x steps / seconds.

Prints one JSON line.  `value` is the headline path's throughput;
`fp64_dynamic` the comparison leg and `speedup_vs_fp64_dynamic` their
ratio; `fp64_fast_dynamic` the same FP64 leg with the fast build of the stage
kernel (FMA contraction and reciprocal quotients, not bit-identical).  `roofline` prices the dominant kernel's algorithmic bytes per
launch against its average launch duration, measured with HIP events on the
context's stream over the timed steps; `mlp_roofline` the learned-closure
kernel's algorithmic FLOPs against the BF16 dense peak (its products are BF16).  `traffic` (HBM bytes per
launch) and the VALU / MFMA busy fractions come from rocprofv3 counter runs
of the same command (profiles/).  `cpu_baseline` times the reference's
own serial path (src/, compiled from its sources into
oracle/_ref/libfslref_dyn.so with real = double and zero-initialised
arrays) on one host core and a bounded sample, or, where that library was
not built, the FP64 CPU oracle (the same algorithm restated in C; also
reported as `cpu_port`) -- a reported baseline, not the target.

    python bench.py [--gpus 1] [--steps 300] [--warmup 10] [--mesh 384x256x256]
                    [--sgs learned|dynamic|constant] [--precision mixed|fp64|fp64fast]
                    [--mach 0.4] [--path fused|unfused] [--model file] [--no-compare] [--no-cpu]
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
BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 MFMA
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


def workload_name(precision, sgs, path, L, H, D, mach, activation=None):
    net = " (%s network)" % {fsl.FSL_MLP_TANH: "tanh", fsl.FSL_MLP_RELU: "ReLU"}[activation] \
        if sgs == "learned" and activation is not None else ""
    return "%s %s-SGS%s %s path, %dx%dx%d cells, Mc %.1f, 3-stage TVD RK" % (
        {"mixed": "FP32+FP64", "fp64": "FP64", "fp64fast": "FP64 (FMA build)"}[precision], sgs, net, path,
        L, H, D, mach)


def spin_up(ctx, cap):
    """Untimed steps until the learned closure has passed a judge and runs on
    the network (the synthetic initial state -- laminar profile plus white
    noise -- may fail the first judge, which leaves the next steps to the
    dynamic fallback).  Returns the number of steps taken; raises when the
    closure is still in fallback after `cap` steps."""
    n = 0
    while True:
        st = ctx.learned_status()
        if st["judged"] >= 1 and not st["in_fallback"] and st["learned_steps"] > 0:
            return n
        if n >= cap:
            raise RuntimeError("learned closure still on the dynamic fallback after %d spin-up steps "
                               "(status %s)" % (n, st))
        ctx.step()
        n += 1


WINDOW_KEYS = ("judged", "fallbacks", "learned_steps", "dynamic_steps")


def time_leg(L, H, D, sgs, precision, path, cfl, mach, steps, warmup, profile=True, spinup_cap=400,
             presteps=0, model=MODEL):
    ctx, case, U0, u_in, _ = shear_case.make(L, H, D, sgs=sgs, precision=precision, fused=(path == "fused"),
                                             cfl=cfl, mach=mach)
    spun = 0
    if sgs == "learned":
        ctx.mlp_load(model)
        ctx.learned_configure(50, 0.10)
        spun = spin_up(ctx, spinup_cap)
    for _ in range(presteps + warmup):
        ctx.step()
    ctx.synchronize()
    ctx.profile_reset()
    ctx.profile(profile)
    ctx.synchronize()
    before = ctx.learned_status() if sgs == "learned" else None
    t0 = time.perf_counter()
    ctx.run(steps)
    ctx.synchronize()
    t = time.perf_counter() - t0
    ctx.profile(False)
    prof = ctx.profile_read()
    sums = ctx.conservation_sums()
    status = window = None
    if sgs == "learned":
        status = ctx.learned_status()
        window = {k: status[k] - before[k] for k in WINDOW_KEYS}
        window["steps"] = steps
        window["in_fallback_at_end"] = status["in_fallback"]
    out = {"ctx": ctx, "case": case, "U0": U0, "u_in": u_in, "t": t, "prof": prof, "sums": sums,
           "status": status, "window": window, "spinup_steps": spun,
           "value": L * H * D * 3 * steps / t / 1e6}
    return out


# Sources whose text determines each profiled kernel's machine code (with
# build.py, which holds the compile flags): a counter figure recorded for
# other sources is stale and is not reported.
KERNEL_SOURCES = {
    ("fused_stage", "mixed"): ["fsl_fused_mixed.hip", "fsl_device_mixed.hpp", "fsl_device.hpp", "fsl_buffer.hpp"],
    ("fused_stage", "fp64"): ["fsl_fused.hip", "fsl_device.hpp", "fsl_buffer.hpp"],
    ("fused_stage", "fp64fast"): ["fsl_fused_fast.hip", "fsl_fused.hip", "fsl_device.hpp", "fsl_buffer.hpp"],
    ("mlp_mu", "mixed"): ["fsl_mlp.hip"],
    ("mlp_mu", "fp64"): ["fsl_mlp.hip"],
}


def sources_sha16(kernel, precision):
    import hashlib
    files = KERNEL_SOURCES.get((kernel, precision))
    if not files:
        return None
    h = hashlib.sha256()
    for f in files + ["../build.py"]:
        with open(os.path.join(ROOT, "freeshearlayer_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def profile_table():
    """Per-kernel figures measured with rocprofv3 counters (profiles/)."""
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    return json.load(open(tpath)).get("entries", {}) if os.path.exists(tpath) else {}


def counter_figures(kernel, precision, workload):
    """(traffic, valu_busy, mfma_busy, provenance) of `kernel` on `workload`
    from profiles/hbm_traffic.json; the figures are None unless they were
    measured on kernels built from the current sources."""
    e = profile_table().get("%s@%s" % (kernel, workload))
    if e is None:
        return None, None, None, {"source": None, "note": "no counter run recorded for this workload"}
    prov = {"source": "profiles/hbm_traffic.json (rocprofv3 --pmc, %s)" % e.get("measured"),
            "sources_sha16": e.get("sources_sha16")}
    cur = sources_sha16(kernel, precision)
    if cur is None or cur != e.get("sources_sha16"):
        prov["note"] = "stale: the kernel sources changed since the counter run (now %s)" % cur
        return None, None, None, prov
    return e.get("traffic"), e.get("valu_busy_pct"), e.get("mfma_busy_pct"), prov


def roofline_of(prof, L, H, D, sgs, precision, workload):
    if not prof:
        return None, {}
    kernels = {k: {"ms_total": round(v[0], 3), "launches": int(v[1]), "avg_us": round(v[0] / v[1] * 1e3, 2)}
               for k, v in sorted(prof.items())}
    name, (ms, cnt) = max(prof.items(), key=lambda kv: kv[1][0])
    b = algorithmic_bytes(name, L, H, D, sgs != "constant")
    if b is None:
        return None, kernels
    avg_s = ms / cnt / 1e3
    traffic, valu, _, prov = counter_figures(name, precision, workload)
    ach = b / avg_s / 1e9
    return ({"bound": "hbm", "kernel": name, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "avg_launch_us": round(avg_s * 1e6, 2),
             "alg_bytes_per_launch": int(b), "valu_busy_pct": valu, "counters": prov}, kernels)


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
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--spinup-cap", type=int, default=400,
                    help="most untimed steps the learned leg may take to leave the dynamic fallback")
    ap.add_argument("--mesh", default="384x256x256")
    ap.add_argument("--sgs", default="learned", choices=["learned", "dynamic", "constant"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64", "fp64fast"])
    ap.add_argument("--path", default="fused", choices=["fused", "unfused"])
    ap.add_argument("--mach", type=float, default=0.4)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--model", default=MODEL, help="learned-closure model file (FSLMLP01 / FSLMLP02)")
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
    act = fsl.mlp_file_read_act(a.model)[2] if a.sgs == "learned" else None
    wl = workload_name(a.precision, a.sgs, a.path, L, H, D, a.mach, act)
    print("[bench] %s: %d warmup + %d timed steps" % (wl, a.warmup, a.steps), file=sys.stderr, flush=True)
    leg = time_leg(L, H, D, a.sgs, a.precision, a.path, a.cfl, a.mach, a.steps, a.warmup, not a.no_profile,
                   spinup_cap=a.spinup_cap, model=a.model)
    ctx = leg["ctx"]
    failures = []
    if not np.all(np.isfinite(leg["sums"])):
        print("bench.py: non-finite state after the timed steps", file=sys.stderr)
        return 1
    roofline, kernels = roofline_of(leg["prof"], L, H, D, a.sgs, a.precision, wl)
    mlp_roofline = None
    if "mlp_mu" in leg["prof"]:
        ms, cnt = leg["prof"]["mlp_mu"]
        tf = L * H * D * MLP_FLOP_PER_CELL / (ms / cnt / 1e3) / 1e12
        mtraffic, valu, mfma, prov = counter_figures("mlp_mu", a.precision, wl)
        mlp_roofline = {"bound": "mfma", "kernel": "mlp_mu", "achieved": round(tf, 2), "peak": BF16_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(tf / BF16_PEAK_TFLOPS, 5),
                        "flop_per_cell": MLP_FLOP_PER_CELL, "mfma_flop_per_cell": MLP_MFMA_FLOP_PER_CELL,
                        "avg_launch_us": round(ms / cnt * 1e3, 2), "launches": int(cnt),
                        "traffic": mtraffic, "alg_bytes_per_launch": 40 * L * H * D,
                        "mfma_busy_pct": mfma, "valu_busy_pct": valu, "counters": prov}
    window = leg["window"]
    if a.sgs == "learned":
        if window["dynamic_steps"] or window["fallbacks"] or window["learned_steps"] != a.steps:
            failures.append("timed steps ran on the dynamic fallback: %s" % window)
        if mlp_roofline is None and not a.no_profile:
            failures.append("no learned-closure kernel launch was timed")
    final_dt = ctx.dt
    ctx.close()
    compare = compare_fast = None
    if not a.no_compare and not (a.sgs == "dynamic" and a.precision == "fp64"):
        print("[bench] comparison leg: FP64 dynamic-SGS, same grid", file=sys.stderr, flush=True)
        # as many untimed steps as the headline leg took, so both time the same flow
        ref = time_leg(L, H, D, "dynamic", "fp64", a.path, a.cfl, a.mach, a.steps, a.warmup, not a.no_profile,
                       presteps=leg["spinup_steps"])
        rl, _ = roofline_of(ref["prof"], L, H, D, "dynamic", "fp64",
                            workload_name("fp64", "dynamic", a.path, L, H, D, a.mach))
        compare = {"value": round(ref["value"], 2), "ms_per_step": round(ref["t"] / a.steps * 1e3, 4),
                   "workload": workload_name("fp64", "dynamic", a.path, L, H, D, a.mach), "roofline": rl}
        ref["ctx"].close()
        if a.path == "fused":
            # the fast build of the FP64 stage kernel (FMA contraction, reciprocal
            # quotients; within 1e-12 of the parity path after one step)
            print("[bench] comparison leg: FP64 (FMA build) dynamic-SGS, same grid", file=sys.stderr, flush=True)
            fast = time_leg(L, H, D, "dynamic", "fp64fast", a.path, a.cfl, a.mach, a.steps, a.warmup,
                            not a.no_profile, presteps=leg["spinup_steps"])
            fwl = workload_name("fp64fast", "dynamic", a.path, L, H, D, a.mach)
            frl, _ = roofline_of(fast["prof"], L, H, D, "dynamic", "fp64fast", fwl)
            compare_fast = {"value": round(fast["value"], 2), "ms_per_step": round(fast["t"] / a.steps * 1e3, 4),
                            "workload": fwl, "roofline": frl}
            fast["ctx"].close()
    cpu = port = None
    if not a.no_cpu:
        print("[bench] CPU reference sample", file=sys.stderr, flush=True)
        cpu = cpu_reference()
        print("[bench] CPU oracle sample", file=sys.stderr, flush=True)
        port = cpu_baseline()
        if cpu is None:
            failures.append("the compiled reference (oracle/_ref/libfslref_dyn.so) is missing: "
                            "cpu_baseline would be the restatement")
            cpu = port
    value = None if failures else round(leg["value"], 2)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(leg["t"] / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision.startswith("fp64") else "f32+f64+bf16" if a.sgs == "learned" else "f32+f64",
        "data": "synthetic shear layer (tanh profile, Mc %.1f, U2/U1 0.5, Re 500, 2%% seeded perturbations, "
                "acoustic CFL %.2f)" % (a.mach, a.cfl),
        "config": {"workload": wl, "mesh": [L, H, D], "stages": 3, "sgs": a.sgs, "precision": a.precision,
                   "path": a.path, "mach": a.mach,
                   "closure": ({"model": os.path.relpath(a.model, ROOT),
                                "activation": {fsl.FSL_MLP_TANH: "tanh", fsl.FSL_MLP_RELU: "relu"}[act],
                                "level": "set per y row by the judge (adapt 2)",
                                "judge_interval": 50, "fallback_tol": 0.10}
                               if a.sgs == "learned" else None)},
        "roofline": roofline,
        "mlp_roofline": mlp_roofline,
        "fp64_dynamic": compare,
        "speedup_vs_fp64_dynamic": (round(leg["value"] / compare["value"], 3)
                                    if compare and value is not None else None),
        "fp64_fast_dynamic": compare_fast,
        "spinup_steps": leg["spinup_steps"],
        "learned_window": window,
        "learned_status": leg["status"],
        "cpu_baseline": cpu,
        "cpu_port": port,
        "kernels": kernels,
        "final_dt": final_dt,
    }
    if failures:
        line["failures"] = failures
        for f in failures:
            print("bench.py: %s" % f, file=sys.stderr)
    print(json.dumps(line))
    return 1 if failures else 0


if __name__ == "__main__":
    sys.exit(main())

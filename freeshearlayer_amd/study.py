"""Side-by-side runs of two paths of the per-step update on the synthetic
shear layer, from one initial state: the FP64 dynamic-Smagorinsky path as
the reference against

  mixed    the mixed-precision path with the same (dynamic) closure
  learned  the learned closure (judge every 50 steps, dynamic fallback
           above 10 % error), in mixed precision unless --precision fp64

Every `every` steps both record delta_theta (mean over the downstream half
and at the outflow plane), delta_w, the integrated resolved TKE and the
mass and energy totals; the result gives their relative differences over
time, the final state's field error, the z energy spectrum of u' on the
centre plane (with --spectra), and each path's Mcell-updates/s per RK stage
(timed without the diagnostics).  With --ref-steps the reference leg stops
early and the comparison covers its steps (spectra at that step).

Usage:
  python -m freeshearlayer_amd.study learned --mesh 384x256x256 --mach 0.4 --steps 2000 --out r.json
  python -m freeshearlayer_amd.study mixed --mesh 384x256x256 --mach 0.3 --steps 2000 --out r.json
  python -m freeshearlayer_amd.study learned --mesh 512x384x384 --mach 0.6 --reynolds 2000 --pert 0.03
      --steps 5000 --ref-steps 1000 --spectra --out r.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import freeshearlayer_amd as fsl
from freeshearlayer_amd import shear_case

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "sgs_mlp.bin")


def log(*a):
    print(*a, flush=True)


def spectrum(ctx, case):
    """z energy spectrum of the fluctuation u' = u - <u>_z on the centre row
    j = HIG/2, averaged over the downstream half of x."""
    S = ctx.download_state()
    L, H, D = case.LEN, case.HIG, case.DEP
    j = H // 2
    u = S[1, L // 2 + 1:L + 1, j, 1:D + 1] / S[0, L // 2 + 1:L + 1, j, 1:D + 1]
    v = S[2, L // 2 + 1:L + 1, j, 1:D + 1] / S[0, L // 2 + 1:L + 1, j, 1:D + 1]
    w = S[3, L // 2 + 1:L + 1, j, 1:D + 1] / S[0, L // 2 + 1:L + 1, j, 1:D + 1]
    E = np.zeros(D // 2 + 1)
    for f in (u, v, w):
        fh = np.fft.rfft(f - f.mean(1, keepdims=True), axis=1) / D
        E += 0.5 * (np.abs(fh) ** 2).mean(0)
    return E, S


def run(path, mesh, args, model):
    L, H, D = mesh
    sgs = "dynamic" if path == "reference" or args.mode == "mixed" else "learned"
    prec = "fp64" if path == "reference" else args.precision
    ctx, case, U, u_in, shear = shear_case.make(L, H, D, sgs=sgs, precision=prec, mach=args.mach,
                                                reynolds=args.reynolds, pert=args.pert, seed=args.seed)
    if sgs == "learned":
        ctx.mlp_load(model)
        ctx.learned_configure(args.interval, args.tol)
        ctx.learned_set_scale(1.0, args.adapt)
    rec = {"path": path, "sgs": sgs, "precision": prec, "step": [], "delta_theta": [], "delta_theta_out": [],
           "delta_w": [], "tke": [], "mass": [], "energy": [], "dt": []}

    def record(step):
        st = shear_case.statistics(ctx, case, u_in, shear.rho0)
        sums = ctx.conservation_sums()
        rec["step"].append(step)
        for k in ("delta_theta", "delta_theta_out", "delta_w", "tke"):
            rec[k].append(st[k])
        rec["mass"].append(float(sums[0]))
        rec["energy"].append(float(sums[4]))
        rec["dt"].append(ctx.dt)

    nsteps = args.ref_steps if (path == "reference" and args.ref_steps) else args.steps
    record(0)
    timed = 0.0
    done = 0
    while done < nsteps:
        n = min(args.every, nsteps - done)
        if args.spectra and args.ref_steps and done < args.ref_steps < done + n:
            n = args.ref_steps - done   # stop on the comparison step
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.run(n)
        ctx.synchronize()
        timed += time.perf_counter() - t0
        done += n
        record(done)
        if args.spectra and done == (args.ref_steps or nsteps):
            rec["spectrum_z"] = spectrum(ctx, case)[0].tolist()
            rec["spectrum_step"] = done
        if not (np.isfinite(rec["mass"][-1]) and np.isfinite(rec["energy"][-1])):
            rec["diverged_at"] = done
            log("  %s: non-finite state after step %d" % (path, done))
            break
        log("  %s %s/%s step %d: delta_theta %.4f, tke %.4e, %.2f ms/step" % (
            path, sgs, prec, done, rec["delta_theta"][-1], rec["tke"][-1], 1e3 * timed / done))
    N = L * H * D
    rec["ms_per_step"] = 1e3 * timed / done
    rec["mcell_updates_per_s_per_stage"] = N * 3 / (timed / done) / 1e6
    if sgs == "learned":
        rec["learned_status"] = ctx.learned_status()
        e, sc = ctx.learned_history()
        rec["judge_error"] = e.tolist()
        rec["judge_scale"] = sc.tolist()
        rec["adapt"] = args.adapt
    S = ctx.download_state() if not args.ref_steps else None
    ctx.close()
    return rec, S


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("mode", choices=["mixed", "learned"])
    ap.add_argument("--mesh", default="384x256x256")
    ap.add_argument("--mach", type=float, default=0.4)
    ap.add_argument("--reynolds", type=float, default=500.0)
    ap.add_argument("--pert", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--model", default=DEFAULT_MODEL)
    ap.add_argument("--interval", type=int, default=50)
    ap.add_argument("--tol", type=float, default=0.10)
    ap.add_argument("--adapt", type=int, default=2, choices=[0, 1, 2, 3],
                    help="level of the network's Cd set by each judge: 0 kept, 1 global, 2 per y row, "
                         "3 per x band and y row")
    ap.add_argument("--spectra", action="store_true")
    ap.add_argument("--ref-steps", type=int, default=0,
                    help="steps of the reference leg when shorter than the candidate's (compared over those)")
    ap.add_argument("--out")
    a = ap.parse_args(argv)
    if a.mode == "mixed":
        a.precision = "mixed"
    mesh = shear_case.parse_mesh(a.mesh)
    log("%s study, mesh %s, Mc %.2f, Re %.0f, %d steps" % (a.mode, a.mesh, a.mach, a.reynolds, a.steps))
    ref, Sa = run("reference", mesh, a, a.model)
    alt, Sb = run("candidate", mesh, a, a.model)

    def rel(k):
        n = min(len(alt[k]), len(ref[k]))   # records at the common steps
        x, y = np.array(alt[k][:n]), np.array(ref[k][:n])
        return (np.abs(x - y) / np.maximum(np.abs(y), 1e-300)).tolist()

    cmp = {k: rel(k) for k in ("delta_theta", "delta_theta_out", "delta_w", "tke", "mass", "energy")}
    field = None
    if Sa is not None and Sb is not None:
        # field error of the final state, per conserved field, against the field's scale
        I = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
        A, B = Sa[I], Sb[I]
        scale = [np.abs(A[0]).max(), np.abs(A[1]).max(), np.abs(A[1]).max(), np.abs(A[1]).max(),
                 np.abs(A[4]).max()]
        field = [float(np.abs(B[f] - A[f]).max() / scale[f]) for f in range(5)]
    out = {"mode": a.mode, "mesh": a.mesh, "mach": a.mach, "reynolds": a.reynolds, "pert": a.pert,
           "steps": a.steps, "reference": ref, "candidate": alt, "relative_difference": cmp,
           "final_field_error": field,
           "speedup": ref["ms_per_step"] / alt["ms_per_step"]}
    if a.spectra:
        Ea, Eb = np.array(ref["spectrum_z"]), np.array(alt["spectrum_z"])
        out["spectrum_rel_difference"] = (np.abs(Eb - Ea) / np.maximum(Ea, 1e-300)).tolist()
    n = min(len(alt["delta_theta"]), len(ref["delta_theta"])) - 1
    log("at step %d: delta_theta %+.2f %%, tke %+.2f %%, mass %.2e, energy %.2e; field error %s; speedup %.2fx "
        "(%.0f vs %.0f Mcell-updates/s per stage)" % (
            alt["step"][n], 100 * (alt["delta_theta"][n] / ref["delta_theta"][n] - 1),
            100 * (alt["tke"][n] / ref["tke"][n] - 1), cmp["mass"][-1], cmp["energy"][-1],
            " ".join("%.1e" % v for v in field) if field else "-", out["speedup"],
            alt["mcell_updates_per_s_per_stage"], ref["mcell_updates_per_s_per_stage"]))
    log("max over time: delta_theta %.2f %%, tke %.2f %%" % (100 * max(cmp["delta_theta"]), 100 * max(cmp["tke"])))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())

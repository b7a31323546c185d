"""Training of the learned SGS closure on this solver's own FP64
dynamic-Smagorinsky runs.

For each convective Mach number and Reynolds number the synthetic shear
layer runs on the fused FP64 dynamic path; after `warmup` steps, every `every` steps the device
sample store receives the features and the dynamic coefficient Cd / 0.15 of
every `stride`-th interior cell (offset cycled so all cells are visited).
The last `holdout` fraction of each run's snapshots is kept for validation.
The network trains with Adam on the pooled samples over a schedule of
learning rates, for each candidate loss weight (|S| Delta / dU)^p (p = 3
weights a sample by its share of the SGS dissipation rho Cd Delta^2 |S|^3)
and initialisation seed; its output layer is then scaled so that the
dissipation-weighted mean coefficient of the training samples matches the
dynamic model's.  Each candidate is reported on the held-out samples and by
the judge (relative SGS-dissipation error against the dynamic model) on each
run's final state, and scored by deployment checks: the closure as
deployed (judge every 50 steps, dynamic fallback above 10 %, the judge's
global rescale) from the initial state at other flow conditions.  The score
is the number of failed judges after the first (the synthetic initial state
is left to the fallback), then their mean error; the best candidate is
written.  The defaults are the recipe of the shipped model.

Options for the early phase of a run (none by default): samples of the
initial state and of every --early-every-th of the first --early steps
(--early-stride, --ic-stride), runs on other meshes (Mc:Re:pert:LxHxD),
deployment checks on other meshes with their own seeds, the first judge
in the score (--score-first), a bound on the judge's profile error
(--profile-tol) and on the final TKE / delta_theta deviation from an FP64
dynamic-model run of each check (--tke-tol).

Usage:
  python -m freeshearlayer_amd.train_sgs --out model.bin [--report r.json]
      [--mesh 192x128x96] [--runs Mc:Re:pert[:mesh],...] [--samples 10000000] ...
"""
import argparse
import json
import sys
import time

import numpy as np

import freeshearlayer_amd as fsl
from freeshearlayer_amd import shear_case


def log(*a):
    print(*a, flush=True)


def generate(mesh, mach, reynolds, pert, warmup, steps, every, stride, seed, early=0, early_stride=None,
             early_every=1, ic_stride=None):
    """One FP64 dynamic-model run; samples every `every` steps after `warmup`,
    and, with early > 0, also of the initial state and of every
    `early_every`-th of the first `early` steps (every `early_stride`-th
    cell): the laminar-plus-noise states and the transition out of them
    that the closure meets at its first judges; `ic_stride` (default
    `early_stride`) is the cell stride on the initial state itself."""
    L, H, D = mesh
    ctx, case, U, u_in, _ = shear_case.make(L, H, D, sgs="dynamic", precision="fp64", mach=mach,
                                            reynolds=reynolds, pert=pert, seed=seed)
    N = L * H * D
    nsnap = steps // every
    per = (N + stride - 1) // stride
    est = early_stride or stride
    ist = ic_stride or est
    nearly = (early // early_every) * ((N + est - 1) // est) + (N + ist - 1) // ist if early > 0 else 0
    ctx.mlp_reserve(nsnap * per + nearly)
    t0 = time.time()
    if early > 0:
        for s in range(early + 1):
            if s:
                ctx.step()
            if s == 0:
                ctx.mlp_collect(ist, offset=0)
            elif s % early_every == 0:
                ctx.mlp_collect(est, offset=(s * 104729) % est)
    for s in range(warmup):
        ctx.step()
        if (s + 1) % 500 == 0:
            log("  Mc %.2f Re %.0f warmup step %d, dt %.3e" % (mach, reynolds, s + 1, ctx.dt))
    n_early = ctx.mlp_sample_count()
    sizes = [n_early] if early > 0 else []
    for q in range(nsnap):
        for _ in range(every):
            ctx.step()
        before = ctx.mlp_sample_count()
        ctx.mlp_collect(stride, offset=(q * 7919) % stride)
        sizes.append(ctx.mlp_sample_count() - before)
    X, t = ctx.mlp_samples()
    st = shear_case.statistics(ctx, case, u_in)
    log("  Mc %.2f Re %.0f pert %.2f: %d steps, %d samples (%d early) in %.1f s; delta_theta %.4g, tke %.4g, "
        "Cd>0 in %.1f %%" % (mach, reynolds, pert, warmup + nsnap * every + early, len(t), n_early,
                            time.time() - t0, st["delta_theta"], st["tke"], 100 * np.mean(t > 0)))
    ctx.mlp_clear_samples()
    return ctx, X, t, np.array(sizes)


def deploy_check(params, norm, act, mesh, mach, reynolds, pert, steps, seed, lcg_seed):
    """The closure as deployed (judge every 50 steps with the dynamic
    fallback and the judge's global rescale) from the initial state at
    another flow condition; returns the judges' level and profile errors,
    the status and the shear-layer statistics of the final state."""
    L, H, D = mesh
    ctx, case, U, u_in, _ = shear_case.make(L, H, D, sgs="learned", precision="mixed", mach=mach,
                                            reynolds=reynolds, pert=pert, seed=seed, lcg_seed=lcg_seed)
    ctx.mlp_set_params(params, norm)
    ctx.mlp_set_activation(act)
    ctx.learned_configure(50, 0.10)
    ctx.learned_set_tolerances(1e30, 1e30)   # scored here, not enforced
    ctx.run(steps)
    err, perr, _, _ = ctx.learned_history(errors=True)
    st = ctx.learned_status()
    stats = shear_case.statistics(ctx, case, u_in)
    ctx.close()
    return [float(e) for e in err], [float(e) for e in perr], st, stats


def reference_statistics(mesh, mach, reynolds, pert, steps, seed, lcg_seed):
    """Shear-layer statistics of the FP64 dynamic-model run of a deployment
    check's case after the same number of steps."""
    L, H, D = mesh
    ctx, case, U, u_in, _ = shear_case.make(L, H, D, sgs="dynamic", precision="fp64", mach=mach,
                                            reynolds=reynolds, pert=pert, seed=seed, lcg_seed=lcg_seed)
    ctx.run(steps)
    stats = shear_case.statistics(ctx, case, u_in)
    ctx.close()
    return {k: float(stats[k]) for k in ("delta_theta", "delta_theta_out", "delta_w", "tke")}


def parse_conditions(spec, mesh):
    """'Mc:Re:pert[:LxHxD],...' -> [(Mc, Re, pert, mesh)], the mesh defaulting to `mesh`."""
    out = []
    for item in spec.split(","):
        v = item.split(":")
        out.append((float(v[0]), float(v[1]), float(v[2]), shear_case.parse_mesh(v[3]) if len(v) > 3 else mesh))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", required=True)
    ap.add_argument("--report")
    ap.add_argument("--mesh", default="192x128x96")
    ap.add_argument("--runs", default="0.3:500:0.02,0.3:2000:0.02,0.5:500:0.02,0.5:2000:0.02",
                    help="Mc:Re:perturbation[:mesh] of the training runs (Re on delta_w0; mesh default --mesh)")
    ap.add_argument("--early", type=int, default=0,
                    help="also sample the initial state and each of the first N steps of every run")
    ap.add_argument("--early-every", type=int, default=1, help="step spacing of the early samples")
    ap.add_argument("--early-stride", type=int, default=0, help="cell stride of the early samples (0: auto)")
    ap.add_argument("--ic-stride", type=int, default=0,
                    help="cell stride on the initial state (0: the early stride)")
    ap.add_argument("--profile-tol", type=float, default=1e30,
                    help="profile-error bound of the judge, used in the deployment score")
    ap.add_argument("--score-first", action="store_true",
                    help="count the first judge (the synthetic initial state) in the deployment score")
    ap.add_argument("--samples", type=int, default=10_000_000, help="total samples (all runs)")
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--steps", type=int, default=3000, help="collection steps per run")
    ap.add_argument("--every", type=int, default=30)
    ap.add_argument("--holdout", type=float, default=0.1)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--schedule", default="8:2e-3,6:1e-3,6:3e-4", help="epochs:lr,...")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--activation", default="tanh", choices=sorted(fsl.MLP_ACTIVATIONS),
                    help="hidden-layer activation of the network")
    ap.add_argument("--weight-power", default="0,3",
                    help="candidate loss weights (|S| Delta / dU)^p")
    ap.add_argument("--seeds", type=int, default=3, help="initialisations per weight power")
    ap.add_argument("--no-calibrate", dest="calibrate", action="store_false")
    ap.add_argument("--eval", default="0.4:500:0.02,0.6:2000:0.02",
                    help="Mc:Re:perturbation[:mesh] of the deployment checks")
    ap.add_argument("--eval-steps", type=int, default=1000, help="steps of each deployment check")
    ap.add_argument("--eval-seed", type=int, default=4242, help="perturbation seed of the deployment checks")
    ap.add_argument("--eval-lcg", type=float, default=0.5772156649,
                    help="disturbance-block generator seed of the deployment checks")
    ap.add_argument("--tke-tol", type=float, default=0.0,
                    help="bound on the final resolved TKE and delta_theta deviation of a deployment check from "
                         "an FP64 dynamic-model run of its case (0: not scored, no reference runs)")
    a = ap.parse_args(argv)

    mesh = shear_case.parse_mesh(a.mesh)
    conds = parse_conditions(a.runs, mesh)
    evals = parse_conditions(a.eval, mesh) if a.eval else []
    nsnap = a.steps // a.every
    n_early = a.early // a.early_every + 1 if a.early else 0
    N0 = mesh[0] * mesh[1] * mesh[2]
    per_run = a.samples / len(conds)
    log("(Mc, Re, pert, mesh) %s: %d snapshots per run, %d early states" % (conds, nsnap, n_early))

    runs, Xtr, ttr, Xva, tva, strides = [], [], [], [], [], []
    for r, (mach, reynolds, pert, m) in enumerate(conds):
        N = m[0] * m[1] * m[2]
        stride = max(1, int(np.ceil(N * nsnap / per_run)))
        # the early states: --early-stride on the --mesh (scaled with the cell
        # count on other meshes), else about a fifth of the run's samples
        if not a.early:
            est = 0
        elif a.early_stride:
            est = max(1, int(round(a.early_stride * N / N0)))
        else:
            est = max(1, int(np.ceil(5 * N * n_early / (N * nsnap / stride))))
        strides.append((stride, est))
        ctx, X, t, sizes = generate(m, mach, reynolds, pert, a.warmup, a.steps, a.every, stride, seed=1000 + r,
                                    early=a.early, early_stride=est, early_every=a.early_every,
                                    ic_stride=max(1, int(round(a.ic_stride * N / N0))) if a.ic_stride else None)
        nva = int(round(a.holdout * len(sizes)))
        cut = int(sizes[:len(sizes) - nva].sum())
        Xtr.append(X[:cut]); ttr.append(t[:cut])
        Xva.append(X[cut:]); tva.append(t[cut:])
        runs.append((mach, reynolds, pert, ctx))
    Xtr = np.concatenate(Xtr); ttr = np.concatenate(ttr)
    log("training samples %d, validation samples %d" % (len(ttr), sum(len(v) for v in tva)))
    # the FP64 dynamic-model statistics each deployment check is compared with
    refs = []
    for mach_e, re_e, pert_e, m in evals:
        if not a.tke_tol:
            refs.append(None)
            continue
        refs.append(reference_statistics(m, mach_e, re_e, pert_e, a.eval_steps, a.eval_seed, a.eval_lcg))
        log("  reference Mc %.2f Re %.0f %s after %d steps: delta_theta %.4g, tke %.4g"
            % (mach_e, re_e, "x".join(map(str, m)), a.eval_steps, refs[-1]["delta_theta"], refs[-1]["tke"]))

    trainer = runs[0][3]
    trainer.mlp_set_activation(a.activation)
    for *_, ctx in runs:
        ctx.mlp_set_activation(a.activation)
    trainer.mlp_load_samples(Xtr, ttr)
    norm0 = trainer.mlp_fit_norm()
    wtr = Xtr[:, 6].astype(np.float64) ** 3   # share of the SGS dissipation rho Cd Delta^2 |S|^3
    report = {"conditions": [(c[0], c[1], c[2], "x".join(map(str, c[3]))) for c in conds],
              "strides": strides, "warmup": a.warmup, "steps": a.steps, "every": a.every, "early": a.early,
              "early_every": a.early_every, "profile_tol": a.profile_tol, "tke_tol": a.tke_tol,
              "train_samples": int(len(ttr)), "batch": a.batch, "activation": a.activation, "schedule": a.schedule,
              "target_variance": float(np.var(ttr)), "eval_references": refs, "candidates": []}
    best = None
    cands = [(float(v), a.seed + k) for v in a.weight_power.split(",") for k in range(a.seeds)]
    for wp, seed in cands:
        trainer.mlp_set_loss_weight(wp)
        params, _ = fsl.mlp_init_params(seed)
        trainer.mlp_set_params(params, norm0)
        losses = []
        t0 = time.time()
        for item in a.schedule.split(","):
            ep, lr = item.split(":")
            ls = trainer.mlp_train(int(ep), a.batch, float(lr), seed=seed * 1000 + len(losses))
            losses += [float(v) for v in ls]
        train_s = time.time() - t0
        params, norm = trainer.mlp_get_params()
        # calibration: scale the output layer so that the dissipation-weighted
        # mean coefficient of the training samples matches the dynamic model's
        ytr = np.clip(trainer.mlp_forward(Xtr), 0.0, 1.0)   # Cd / 0.15 after the clip
        alpha = float((wtr * ttr).sum() / max((wtr * ytr).sum(), 1e-30)) if a.calibrate else 1.0
        params = params.copy()
        params[4736:4801] *= alpha
        cand = {"weight_power": wp, "seed": seed, "epoch_loss": losses, "train_seconds": train_s, "alpha": alpha,
                "runs": []}
        for (mach, reynolds, pert, ctx), Xv, tv in zip(runs, Xva, tva):
            ctx.mlp_set_params(params, norm)
            yv = np.clip(ctx.mlp_forward(Xv), 0.0, 1.0)
            wv = Xv[:, 6].astype(np.float64) ** 3
            mse = float(np.mean((yv - tv) ** 2))
            bias = float((wv * (yv - tv)).sum() / max((wv * tv).sum(), 1e-30))
            ctx.ghost_fill()
            err, perr, _, dd, dl = ctx.learned_judge(errors=True)
            cand["runs"].append({"mach": mach, "reynolds": reynolds, "pert": pert, "val_samples": int(len(tv)),
                                 "val_mse": mse, "val_r2": 1 - mse / max(float(np.var(tv)), 1e-30),
                                 "val_dissipation_bias": bias, "judge_error": err, "judge_profile_error": perr,
                                 "dissipation_dynamic": dd, "dissipation_learned": dl})
        cand["deploy"] = []
        fails, errs = 0, []
        for (mach_e, re_e, pert_e, m), ref in zip(evals, refs):
            dep, pdep, dst, stats = deploy_check(params, norm, a.activation, m, mach_e, re_e, pert_e, a.eval_steps,
                                                 a.eval_seed, a.eval_lcg)
            dtke = stats["tke"] / ref["tke"] - 1 if ref else 0.0
            dth = stats["delta_theta"] / ref["delta_theta"] - 1 if ref else 0.0
            cand["deploy"].append({"mach": mach_e, "reynolds": re_e, "pert": pert_e, "mesh": "x".join(map(str, m)),
                                   "every": 50, "judge_error": dep, "judge_profile_error": pdep, "status": dst,
                                   "tke_rel": dtke, "delta_theta_rel": dth})
            # a judge fails on the level error (0.10) or the profile error
            # (profile_tol); the first (synthetic initial state) only with
            # --score-first; with --tke-tol the final resolved TKE and
            # delta_theta fail beyond that fraction of the dynamic model's
            first = 0 if a.score_first else 1
            fails += sum(e > 0.10 or p > a.profile_tol for e, p in zip(dep[first:], pdep[first:]))
            if ref:
                fails += int(abs(dtke) > a.tke_tol) + int(abs(dth) > a.tke_tol)
            errs += [e + (p if a.profile_tol < 1e30 else 0.0) for e, p in zip(dep[first:], pdep[first:])]
        score = (fails, float(np.mean(errs)) if errs else 0.0)
        cand["score"] = list(score)
        log("  weight power %.1f seed %d: %.1f s, final loss %.3e, alpha %.3f; %s" % (
            wp, seed, train_s, losses[-1], alpha, "; ".join(
                "Mc %.2f Re %.0f R^2 %.3f judge %.3f/%.3f" % (r["mach"], r["reynolds"], r["val_r2"], r["judge_error"],
                                                          r["judge_profile_error"])
                for r in cand["runs"])))
        for d in cand["deploy"]:
            log("    deployment Mc %.2f Re %.0f %s: tke %+.2f %%, delta_theta %+.2f %%; %d of %d judges failed; "
                "errors %s; profile errors %s" % (
                    d["mach"], d["reynolds"], d["mesh"], 100 * d["tke_rel"], 100 * d["delta_theta_rel"],
                    sum(e > 0.10 or p > a.profile_tol for e, p in zip(d["judge_error"], d["judge_profile_error"])),
                    len(d["judge_error"]), " ".join("%.3f" % e for e in d["judge_error"]),
                    " ".join("%.3f" % e for e in d["judge_profile_error"])))
        report["candidates"].append(cand)
        if best is None or score < best[0]:
            best = (score, wp, seed, params, norm)
    trainer.mlp_clear_samples()
    score, wp, seed, params, norm = best
    report["selected"] = {"weight_power": wp, "seed": seed, "score": list(score)}
    log("selected weight power %.1f seed %d (%d failed judges, mean error %.3f)" % (wp, seed, score[0], score[1]))
    fsl.mlp_file_write_act(a.out, params, norm, a.activation)
    log("wrote", a.out)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    for *_, ctx in runs:
        ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

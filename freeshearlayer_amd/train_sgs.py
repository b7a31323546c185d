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
written.

Usage:
  python -m freeshearlayer_amd.train_sgs --out model.bin [--report r.json]
      [--mesh 192x128x96] [--mach 0.3,0.5] [--reynolds 500,2000] [--samples 10000000] ...
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


def generate(mesh, mach, reynolds, warmup, steps, every, stride, seed):
    L, H, D = mesh
    ctx, case, U, u_in, _ = shear_case.make(L, H, D, sgs="dynamic", precision="fp64", mach=mach,
                                            reynolds=reynolds, seed=seed)
    N = L * H * D
    nsnap = steps // every
    per = (N + stride - 1) // stride
    ctx.mlp_reserve(nsnap * per)
    t0 = time.time()
    for s in range(warmup):
        ctx.step()
        if (s + 1) % 500 == 0:
            log("  Mc %.2f Re %.0f warmup step %d, dt %.3e" % (mach, reynolds, s + 1, ctx.dt))
    sizes = []
    for q in range(nsnap):
        for _ in range(every):
            ctx.step()
        before = ctx.mlp_sample_count()
        ctx.mlp_collect(stride, offset=(q * 7919) % stride)
        sizes.append(ctx.mlp_sample_count() - before)
    X, t = ctx.mlp_samples()
    st = shear_case.statistics(ctx, case, u_in)
    log("  Mc %.2f Re %.0f: %d steps, %d samples in %.1f s; delta_theta %.4g, tke %.4g, Cd>0 in %.1f %%"
        % (mach, reynolds, warmup + nsnap * every, len(t), time.time() - t0, st["delta_theta"], st["tke"],
           100 * np.mean(t > 0)))
    ctx.mlp_clear_samples()
    return ctx, X, t, np.array(sizes)


def deploy_check(params, norm, mesh, mach, reynolds, steps, seed):
    """The closure as deployed (judge every 50 steps with the dynamic
    fallback and the judge's global rescale) from the initial state at
    another flow condition; returns the judge errors and the status."""
    L, H, D = mesh
    ctx, case, U, u_in, _ = shear_case.make(L, H, D, sgs="learned", precision="mixed", mach=mach,
                                            reynolds=reynolds, seed=seed)
    ctx.mlp_set_params(params, norm)
    ctx.learned_configure(50, 0.10)
    ctx.run(steps)
    err, _ = ctx.learned_history()
    st = ctx.learned_status()
    ctx.close()
    return [float(e) for e in err], st


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", required=True)
    ap.add_argument("--report")
    ap.add_argument("--mesh", default="192x128x96")
    ap.add_argument("--mach", default="0.3,0.5")
    ap.add_argument("--reynolds", default="500,2000", help="Re on delta_w0 of the training runs")
    ap.add_argument("--samples", type=int, default=10_000_000, help="total samples (all runs)")
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--steps", type=int, default=3000, help="collection steps per run")
    ap.add_argument("--every", type=int, default=30)
    ap.add_argument("--holdout", type=float, default=0.1)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--schedule", default="8:2e-3,6:1e-3,6:3e-4", help="epochs:lr,...")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--weight-power", default="0,3",
                    help="candidate loss weights (|S| Delta / dU)^p")
    ap.add_argument("--seeds", type=int, default=3, help="initialisations per weight power")
    ap.add_argument("--no-calibrate", dest="calibrate", action="store_false")
    ap.add_argument("--eval", default="0.4:500,0.6:2000", help="Mc:Re of the deployment checks")
    ap.add_argument("--eval-steps", type=int, default=1000, help="steps of each deployment check")
    a = ap.parse_args(argv)

    mesh = shear_case.parse_mesh(a.mesh)
    conds = [(float(m), float(re)) for m in a.mach.split(",") for re in a.reynolds.split(",")]
    evals = [tuple(float(v) for v in e.split(":")) for e in a.eval.split(",")] if a.eval else []
    N = mesh[0] * mesh[1] * mesh[2]
    nsnap = a.steps // a.every
    stride = max(1, int(np.ceil(N * nsnap * len(conds) / a.samples)))
    log("mesh %s, (Mc, Re) %s: %d snapshots per run, stride %d" % (a.mesh, conds, nsnap, stride))

    runs, Xtr, ttr, Xva, tva = [], [], [], [], []
    for r, (mach, reynolds) in enumerate(conds):
        ctx, X, t, sizes = generate(mesh, mach, reynolds, a.warmup, a.steps, a.every, stride, seed=1000 + r)
        nva = int(round(a.holdout * len(sizes)))
        cut = int(sizes[:len(sizes) - nva].sum())
        Xtr.append(X[:cut]); ttr.append(t[:cut])
        Xva.append(X[cut:]); tva.append(t[cut:])
        runs.append((mach, reynolds, ctx))
    Xtr = np.concatenate(Xtr); ttr = np.concatenate(ttr)
    log("training samples %d, validation samples %d" % (len(ttr), sum(len(v) for v in tva)))

    trainer = runs[0][2]
    trainer.mlp_load_samples(Xtr, ttr)
    norm0 = trainer.mlp_fit_norm()
    wtr = Xtr[:, 6].astype(np.float64) ** 3   # share of the SGS dissipation rho Cd Delta^2 |S|^3
    report = {"mesh": a.mesh, "conditions": conds, "stride": stride, "warmup": a.warmup, "steps": a.steps,
              "every": a.every, "train_samples": int(len(ttr)), "batch": a.batch, "schedule": a.schedule,
              "target_variance": float(np.var(ttr)), "candidates": []}
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
        for (mach, reynolds, ctx), Xv, tv in zip(runs, Xva, tva):
            ctx.mlp_set_params(params, norm)
            yv = np.clip(ctx.mlp_forward(Xv), 0.0, 1.0)
            wv = Xv[:, 6].astype(np.float64) ** 3
            mse = float(np.mean((yv - tv) ** 2))
            bias = float((wv * (yv - tv)).sum() / max((wv * tv).sum(), 1e-30))
            ctx.ghost_fill()
            err, dd, dl = ctx.learned_judge()
            cand["runs"].append({"mach": mach, "reynolds": reynolds, "val_samples": int(len(tv)), "val_mse": mse,
                                 "val_r2": 1 - mse / max(float(np.var(tv)), 1e-30),
                                 "val_dissipation_bias": bias, "judge_error": err,
                                 "dissipation_dynamic": dd, "dissipation_learned": dl})
        cand["deploy"] = []
        fails, errs = 0, []
        for mach_e, re_e in evals:
            dep, dst = deploy_check(params, norm, mesh, mach_e, re_e, a.eval_steps, seed=4242)
            cand["deploy"].append({"mach": mach_e, "reynolds": re_e, "every": 50, "judge_error": dep, "status": dst})
            # the synthetic initial state is left to the fallback: its judge is not scored
            fails += sum(e > 0.10 for e in dep[1:])
            errs += dep[1:]
        score = (fails, float(np.mean(errs)) if errs else 0.0)
        cand["score"] = list(score)
        log("  weight power %.1f seed %d: %.1f s, final loss %.3e, alpha %.3f; %s" % (
            wp, seed, train_s, losses[-1], alpha, "; ".join(
                "Mc %.2f Re %.0f R^2 %.3f judge %.3f" % (r["mach"], r["reynolds"], r["val_r2"], r["judge_error"])
                for r in cand["runs"])))
        for d in cand["deploy"]:
            log("    deployment Mc %.2f Re %.0f: %d of %d later judges failed; errors %s" % (
                d["mach"], d["reynolds"], sum(e > 0.10 for e in d["judge_error"][1:]), len(d["judge_error"]) - 1,
                " ".join("%.3f" % e for e in d["judge_error"])))
        report["candidates"].append(cand)
        if best is None or score < best[0]:
            best = (score, wp, seed, params, norm)
    trainer.mlp_clear_samples()
    score, wp, seed, params, norm = best
    report["selected"] = {"weight_power": wp, "seed": seed, "score": list(score)}
    log("selected weight power %.1f seed %d (%d failed later judges, mean error %.3f)" % (wp, seed, score[0], score[1]))
    fsl.mlp_file_write(a.out, params, norm)
    log("wrote", a.out)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    for _, _, ctx in runs:
        ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""The synthetic spatially developing shear layer used by the benchmark, the
training of the learned closure and the precision / closure comparisons:
the reference case's items (layer2.ini spacing, Pr, Cs, 3-stage TVD RK)
with the mesh replaced, a tanh inflow profile U(y) = Um + dU/2 tanh(2y /
delta_w0) of velocity ratio U2/U1 and convective Mach number Mc = dU / 2c,
uniform p0 and rho0, mu_L from the Reynolds number on delta_w0, seeded
perturbations in |y| < delta_w0, the reference's inflow disturbance block,
and the acoustic CFL time step."""
import numpy as np

import freeshearlayer_amd as fsl

SGS = {"constant": fsl.FSL_SGS_CONSTANT, "dynamic": fsl.FSL_SGS_DYNAMIC, "learned": fsl.FSL_SGS_LEARNED}
PRECISION = {"fp64": fsl.FSL_PREC_FP64, "mixed": fsl.FSL_PREC_MIXED, "fp64fast": fsl.FSL_PREC_FP64_FAST}


def parse_mesh(s):
    L, H, D = (int(v) for v in s.lower().split("x"))
    return L, H, D


def initial(L, H, D, mach=0.3, ratio=0.5, reynolds=500.0, pert=0.02, seed=None):
    """Case items, initial state and inflow profile (host); returns (case, U, u_in, shear)."""
    case = fsl.case_reference(L, H, D)
    shear = fsl.shear_defaults()
    shear.mach_c = mach
    shear.ratio = ratio
    shear.reynolds = reynolds
    shear.pert = pert
    if seed is not None:
        shear.seed = seed
    U, u_in, mu_L = fsl.ic_shear_layer(case, shear)
    case.mu_L = mu_L
    case.Pr_L = 0.72
    case.nStages = 3
    case.BL_HIG = min(5, H - H // 2)
    return case, U, u_in, shear


def make(L, H, D, sgs="dynamic", precision="fp64", fused=True, cfl=0.5, mach=0.3, ratio=0.5,
         reynolds=500.0, pert=0.02, seed=None, lcg_seed=0.317391239817):
    """Context holding the initial state; returns (ctx, case, U, u_in, shear)."""
    case, U, u_in, shear = initial(L, H, D, mach, ratio, reynolds, pert, seed)
    cfg = fsl.config(sgs_model=SGS[sgs], precision=PRECISION[precision], cfl_mode=fsl.FSL_CFL_ACOUSTIC,
                     cfl=cfl, fused=1 if fused else 0, rho_in=shear.rho0, p_in=shear.p0, p_out=shear.p0)
    ctx = fsl.Context(case, cfg)
    ctx.set_inflow(u_in)
    ctx.upload_state(U)
    ctx.set_lcg(lcg_seed, 0)
    return ctx, case, U, u_in, shear


def streams(u_in):
    """(U1, U2): the fast and the slow stream of an inflow profile."""
    return float(np.max(u_in)), float(np.min(u_in))


def statistics(ctx, case, u_in, rho0=1.0):
    """delta_theta, delta_w (means over the downstream half), delta_theta at
    the outflow plane and the integrated resolved TKE of the step state."""
    U1, U2 = streams(u_in)
    return fsl.shear_statistics(case, ctx.plane_profiles(), U1, U2, rho0)

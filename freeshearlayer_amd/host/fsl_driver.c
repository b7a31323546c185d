/*
 * fsl_driver.c -- the time-step driver (ANSI C) over the device operations
 * of include/fsl.h, in the order of the reference's main loop
 * (src/layer2.c:141-185): CFL rescale, disturbance-block draw, then the RK
 * stages.  Probes and output stay with the caller.
 */
#include "fsl.h"

int fsl_step(fsl_ctx *ctx)
{
    fsl_case c;
    fsl_block b;
    int rc, s;
    fsl_get_case(ctx, &c);
    rc = fsl_check_co_num(ctx, (double *)0);
    if (rc != FSL_OK) return rc;
    fsl_set_step(ctx, fsl_get_step(ctx) + 1);
    fsl_get_block_state(ctx, &b);
    fsl_block_advance(&b, &c);
    fsl_set_block_state(ctx, &b);
    for (s = 1; s <= c.nStages; s++) {
        rc = fsl_stage(ctx, s);
        if (rc != FSL_OK) return rc;
    }
    return FSL_OK;
}

int fsl_run(fsl_ctx *ctx, int nsteps)
{
    int n, rc;
    for (n = 0; n < nsteps; n++) {
        rc = fsl_step(ctx);
        if (rc != FSL_OK) return rc;
    }
    return FSL_OK;
}

/*
 * fsl_layer2_main.c -- layer2-compatible driver (ANSI C) on the GPU.
 *
 * Behaves like the reference's serial executable (src/layer2.c:117-195):
 * reads layer2.ini (or an mpi_layer2.bin), the LCG seed from stdin, starts
 * from the two-stream initial condition or from backup.dat (Answer = 1),
 * advances numstep steps, stops early on SIGINT or on a stopfile holding 's'
 * (src-par/mpi_layer2.c:169-191), appends the probes record of every step to
 * probes.dat (src/probes.c), writes the Tecplot frame <step>.plt every f_step
 * steps and at the end (src/output.c), and writes backup.dat (src/finalize.c)
 * -- the frame and backup.dat not after a failed step.  stdout is
 * line-buffered so a supervisor sees each step's line.
 *
 *   fsl_layer2 [-i layer2.ini | -b mpi_layer2.bin] [-n steps] [--const-cs]
 *              [--fused] [--mixed] [--acoustic CFL] [--learned MODEL] [--real4]
 *              [--no-output] [-q] < seed
 *
 * --real4 writes probes.dat and backup.dat with 4-byte reals (the layout of
 * the shipped float build, which the Favre post-processor reads);
 * --no-output skips probes.dat and the frames.
 *
 * --learned MODEL runs the learned SGS closure (parameter file as written by
 * fsl_mlp_file_write or fsl_mlp_file_write_act, with its activation),
 * judged against the dynamic model every 50 steps.
 */
#include "fsl.h"

#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static volatile sig_atomic_t cont_flag = 1;

static void on_sigint(int s)
{
    (void)s;
    cont_flag = 0;
}

static int stopfile_says_stop(void)
{
    FILE *f = fopen("stopfile", "r");
    int ch;
    if (!f) return 0;
    ch = fgetc(f);
    fclose(f);
    return ch == 's';
}

int main(int argc, char **argv)
{
    const char *ini = "layer2.ini", *bin = 0, *model = 0;
    long nsteps = -1;
    int quiet = 0, real_bytes = 8, output = 1, a, rc;
    double X = 0.0, total = 0.0;
    fsl_case c;
    fsl_config cfg;
    fsl_ctx *ctx = 0;
    double *U;
    size_t G;

    setvbuf(stdout, 0, _IOLBF, 0);
    fsl_config_defaults(&cfg);
    for (a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "-i") && a + 1 < argc) ini = argv[++a];
        else if (!strcmp(argv[a], "-b") && a + 1 < argc) bin = argv[++a];
        else if (!strcmp(argv[a], "-n") && a + 1 < argc) nsteps = atol(argv[++a]);
        else if (!strcmp(argv[a], "--const-cs")) cfg.sgs_model = FSL_SGS_CONSTANT;
        else if (!strcmp(argv[a], "--fused")) cfg.fused = 1;
        else if (!strcmp(argv[a], "--mixed")) cfg.precision = FSL_PREC_MIXED;
        else if (!strcmp(argv[a], "--acoustic") && a + 1 < argc) {
            cfg.cfl_mode = FSL_CFL_ACOUSTIC;
            cfg.cfl = atof(argv[++a]);
        } else if (!strcmp(argv[a], "--learned") && a + 1 < argc) {
            cfg.sgs_model = FSL_SGS_LEARNED;
            model = argv[++a];
        } else if (!strcmp(argv[a], "--real4")) real_bytes = 4;
        else if (!strcmp(argv[a], "--no-output")) output = 0;
        else if (!strcmp(argv[a], "-q")) quiet = 1;
        else {
            fprintf(stderr, "usage: %s [-i ini | -b bin] [-n steps] [--const-cs] [--fused] [--mixed] "
                            "[--acoustic CFL] [--learned MODEL] [--real4] [--no-output] [-q] < seed\n", argv[0]);
            return 2;
        }
    }
    signal(SIGINT, on_sigint);
    if (scanf("%lf", &X) == 1) {
        if (!quiet) printf("Random number [0...1]: X = %f\n", X);
    } else {
        printf("Failed to read random number.\n");
    }
    rc = bin ? fsl_case_read_bin(bin, &c) : fsl_case_read_ini(ini, &c);
    if (rc != FSL_OK) {
        fprintf(stderr, "cannot read case file %s\n", bin ? bin : ini);
        return 1;
    }
    if (nsteps < 0) nsteps = (long)c.numstep;
    if (cfg.fused && c.nStages != 3) {
        fprintf(stderr, "the fused path runs 3 stages; nStages = %d\n", c.nStages);
        return 1;
    }
    G = (size_t)(c.LEN + 2) * (c.HIG + 2) * (c.DEP + 2);
    U = (double *)calloc(5 * G, sizeof(double));
    if (!U) return 1;
    {
        unsigned step0 = 0;
        double dt = c.deltaT;
        if (c.Answer) {
            if (fsl_backup_read("backup.dat", &c, U, &dt, &step0, 0, real_bytes) != FSL_OK) {
                puts("Cannot open \"backup.dat\" file to read");
                return 1;
            }
        } else {
            fsl_ic_reference(&c, U);
        }
        rc = fsl_create(&c, &cfg, &ctx);
        if (rc != FSL_OK) {
            fprintf(stderr, "fsl_create failed (%d)\n", rc);
            return 1;
        }
        fsl_set_dt(ctx, dt);
        fsl_set_step(ctx, step0);
    }
    if (model) {
        static float params[FSL_MLP_NPARAM];
        float norm[16];
        int act = FSL_MLP_TANH;
        if (fsl_mlp_file_read_act(model, params, norm, &act) != FSL_OK || fsl_mlp_set_params(ctx, params, norm) != FSL_OK ||
            fsl_mlp_set_activation(ctx, act) != FSL_OK) {
            fprintf(stderr, "cannot load the learned closure from %s\n", model);
            fsl_destroy(ctx);
            free(U);
            return 1;
        }
    }
    fsl_set_lcg(ctx, X, 0);
    rc = fsl_upload_state(ctx, 0, U);
    while (rc == FSL_OK && cont_flag && (long)fsl_get_step(ctx) < nsteps) {
        rc = fsl_step(ctx);
        total += fsl_get_dt(ctx);
        if (!quiet) printf("Timestep no.: %u, total time: %f sec.\n", fsl_get_step(ctx), total);
        if (rc == FSL_OK && output) {   /* src/layer2.c:180-183 */
            rc = fsl_write_probes(ctx, "probes.dat", real_bytes);
            if (rc == FSL_OK && c.f_step && fsl_get_step(ctx) % c.f_step == 0) rc = fsl_write_plt(ctx, 0);
        }
        if (stopfile_says_stop()) break;
    }
    if (rc == FSL_OK && output) rc = fsl_write_plt(ctx, 0);   /* src/layer2.c:187-188 */
    if (rc != FSL_OK) fprintf(stderr, "step failed: %s\n", fsl_last_error(ctx));
    if (model && !quiet) {
        fsl_learned_status st;
        fsl_learned_get_status(ctx, &st);
        printf("Learned closure: %lld steps on the network, %lld on the dynamic model, %lld judges, "
               "%lld fallbacks\n", st.learned_steps, st.dynamic_steps, st.judged, st.fallbacks);
    }
    /* src/finalize.c:14-34 writes backup.dat at the end of the run.  A run
       that ended on a failed step (e.g. a non-finite CFL reduction) leaves
       the previous backup.dat -- with Answer = 1 the run's own input --
       untouched instead of overwriting it with the diverged state. */
    if (rc == FSL_OK && fsl_download_state(ctx, 0, U) == FSL_OK)
        fsl_backup_write("backup.dat", &c, U, fsl_get_dt(ctx), fsl_get_step(ctx), real_bytes);
    else if (rc != FSL_OK)
        fprintf(stderr, "backup.dat not written: the last step failed\n");
    fsl_destroy(ctx);
    free(U);
    puts("Simulation finished!");
    return rc == FSL_OK ? 0 : 1;
}

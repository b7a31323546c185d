/*
 * layer2_dropin.c -- a host program written against the reference's own
 * entry-point names (include/fsl_layer2.h, libfsl_layer2.so), the way the
 * reference's main drives its hot path (src/layer2.c:117-195): real***
 * arrays from Array3D, checkCoNum, the disturbance-block draw with Random,
 * then per stage BounCondInGhostCells / Reconstruction / BounCondOnInterfaces
 * / Fluxes / Evolution.  The three calls the drop-in adds (attach, set_block,
 * sync_host) sit where the integration guide puts them.  Test program only:
 * the parity tests compare what it writes with the oracle.
 *
 *   layer2_dropin run    L H D dynamic nStages nsteps X0 state.bin out.bin
 *   layer2_dropin filter L H D field.bin out.bin
 *   layer2_dropin dynarg L H D state.bin out.bin
 *
 * state.bin: 5 dense cell arrays (L+2)(H+2)(D+2), k fastest.  out.bin of
 * "run": the 5 arrays, mu_SGS, then deltaT, X, Ud, Vd, Wd (double), counter,
 * k_min, k_max, status (int32).  "filter": the filtered array.  "dynarg":
 * two int32 statuses.
 */
#include "fsl_layer2.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static unsigned LEN, HIG, DEP;

static double ***U1, ***U2, ***U3, ***U4, ***U5, ***mu_SGS;

static void load_cells(FILE *f, double ***a)
{
    unsigned i, j;
    for (i = 0; i < LEN + 2; i++)
        for (j = 0; j < HIG + 2; j++)
            if (fread(a[i][j], sizeof(double), DEP + 2, f) != DEP + 2) {
                fprintf(stderr, "short state file\n");
                exit(2);
            }
}

static void save_cells(FILE *f, double ***a)
{
    unsigned i, j;
    for (i = 0; i < LEN + 2; i++)
        for (j = 0; j < HIG + 2; j++) fwrite(a[i][j], sizeof(double), DEP + 2, f);
}

/* case items of the reference's layer2.ini with the mesh replaced */
static fsl_case reference_case(int nStages)
{
    fsl_case c;
    memset(&c, 0, sizeof c);
    c.LEN = LEN; c.HIG = HIG; c.DEP = DEP;
    c.deltaX = 1.6666667; c.deltaY = 1.; c.deltaZ = 1.5; c.deltaT = 0.0007;
    c.Answer = 0; c.numstep = 3000;
    c.mu_L = 0.253358; c.Pr_L = 0.72; c.Pr_T = 0.8; c.Cs = 0.01;
    c.BL_HIG = (HIG - HIG / 2 < 5) ? HIG - HIG / 2 : 5;
    c.Ua = 35.; c.Va = 35.; c.Wa = 35.;
    c.Ns_min = 10; c.Nst = 50; c.f_step = 100000;
    c.nStages = nStages; c.maxCoNum = 0.1;
    return c;
}

static void alloc_state(void)
{
    U1 = Array3D(LEN + 2, HIG + 2, DEP + 2);
    U2 = Array3D(LEN + 2, HIG + 2, DEP + 2);
    U3 = Array3D(LEN + 2, HIG + 2, DEP + 2);
    U4 = Array3D(LEN + 2, HIG + 2, DEP + 2);
    U5 = Array3D(LEN + 2, HIG + 2, DEP + 2);
    mu_SGS = Array3D(LEN + 2, HIG + 2, DEP + 2);
}

static int run(int argc, char **argv)
{
    int dynamic, nStages, nsteps, Stage, counter = 0, st;
    unsigned k_min = 0, k_max = 0, step = 0;
    double X, Ud = 0, Vd = 0, Wd = 0, deltaT;
    fsl_case c;
    fsl_config cfg;
    FILE *f;
    if (argc != 11) return 2;
    dynamic = atoi(argv[5]);
    nStages = atoi(argv[6]);
    nsteps = atoi(argv[7]);
    X = atof(argv[8]);
    c = reference_case(nStages);
    alloc_state();
    f = fopen(argv[9], "rb");
    if (!f) return 2;
    load_cells(f, U1); load_cells(f, U2); load_cells(f, U3); load_cells(f, U4); load_cells(f, U5);
    fclose(f);
    deltaT = c.deltaT;

    fsl_config_defaults(&cfg);                  /* FP64, unfused, convective CFL, as shipped */
    cfg.sgs_model = dynamic ? FSL_SGS_DYNAMIC : FSL_SGS_CONSTANT;
    if (fsl_layer2_attach(&c, &cfg, U1, U2, U3, U4, U5, mu_SGS, &deltaT) != FSL_OK) {
        fprintf(stderr, "attach failed: %s\n", fsl_layer2_error());
        return 1;
    }
    while (step < (unsigned)nsteps) {
        checkCoNum();
        ++step;
        if (counter > 0) counter--;             /* src/layer2.c:152-162 */
        else {
            counter = (int)(c.Ns_min + c.Nst * (X = Random(X)));
            k_min = (unsigned)(DEP * (X = Random(X)));
            k_max = (unsigned)(k_min + DEP * (X = Random(X)));
            if (k_max > DEP) k_max -= DEP;
            X = Random(X); Ud = c.Ua * (1 - X - X);
            X = Random(X); Vd = c.Va * (1 - X - X);
            X = Random(X); Wd = c.Wa * (1 - X - X);
        }
        fsl_layer2_set_block(k_min, k_max, Ud, Vd, Wd);
        for (Stage = 1; Stage <= nStages; Stage++) {
            BounCondInGhostCells();
            Reconstruction();
            BounCondOnInterfaces();
            Fluxes();
            Evolution(nStages, Stage);
        }
        fsl_layer2_sync_host();                 /* before Probes() / Output() */
    }
    st = fsl_layer2_status();
    if (st != FSL_OK) fprintf(stderr, "drop-in error %d: %s\n", st, fsl_layer2_error());
    f = fopen(argv[10], "wb");
    if (!f) return 2;
    save_cells(f, U1); save_cells(f, U2); save_cells(f, U3); save_cells(f, U4); save_cells(f, U5);
    save_cells(f, mu_SGS);
    fwrite(&deltaT, sizeof deltaT, 1, f);
    fwrite(&X, sizeof X, 1, f);
    fwrite(&Ud, sizeof Ud, 1, f); fwrite(&Vd, sizeof Vd, 1, f); fwrite(&Wd, sizeof Wd, 1, f);
    fwrite(&counter, sizeof counter, 1, f);
    fwrite(&k_min, sizeof k_min, 1, f); fwrite(&k_max, sizeof k_max, 1, f);
    fwrite(&st, sizeof st, 1, f);
    fclose(f);
    fsl_layer2_detach();
    free3D(U1, LEN + 2, HIG + 2); free3D(U2, LEN + 2, HIG + 2); free3D(U3, LEN + 2, HIG + 2);
    free3D(U4, LEN + 2, HIG + 2); free3D(U5, LEN + 2, HIG + 2); free3D(mu_SGS, LEN + 2, HIG + 2);
    return st == FSL_OK ? 0 : 1;
}

static int filter(int argc, char **argv)
{
    double h[3] = {1. / 3., 1. / 3., 1. / 3.};
    double ***a, ***r;
    fsl_case c;
    fsl_config cfg;
    FILE *f;
    if (argc != 7) return 2;
    c = reference_case(3);
    alloc_state();
    a = Array3D(LEN + 2, HIG + 2, DEP + 2);
    f = fopen(argv[5], "rb");
    if (!f) return 2;
    load_cells(f, a);
    fclose(f);
    fsl_config_defaults(&cfg);
    if (fsl_layer2_attach(&c, &cfg, U1, U2, U3, U4, U5, mu_SGS, 0) != FSL_OK) return 1;
    r = filter_(a, h);                           /* src/turbulence.c:17: a new array */
    if (!r || fsl_layer2_status() != FSL_OK) return 1;
    f = fopen(argv[6], "wb");
    if (!f) return 2;
    save_cells(f, r);
    fclose(f);
    free3D(r, LEN + 2, HIG + 2);
    free3D(a, LEN + 2, HIG + 2);
    fsl_layer2_detach();
    return 0;
}

/* DynamicSmagorinsky keeps its signature; the arrays must be the attached
   ones (src/fluxes.c:53 passes U1..U4, mu_SGS) */
static int dynarg(int argc, char **argv)
{
    fsl_case c;
    fsl_config cfg;
    int st[2];
    FILE *f;
    if (argc != 7) return 2;
    c = reference_case(3);
    alloc_state();
    f = fopen(argv[5], "rb");
    if (!f) return 2;
    load_cells(f, U1); load_cells(f, U2); load_cells(f, U3); load_cells(f, U4); load_cells(f, U5);
    fclose(f);
    fsl_config_defaults(&cfg);
    if (fsl_layer2_attach(&c, &cfg, U1, U2, U3, U4, U5, mu_SGS, 0) != FSL_OK) return 1;
    DynamicSmagorinsky(U1, U2, U3, U4, mu_SGS);
    fsl_layer2_sync_host();
    st[0] = fsl_layer2_status();
    DynamicSmagorinsky(U2, U1, U3, U4, mu_SGS);  /* not the attached order */
    st[1] = fsl_layer2_status();
    f = fopen(argv[6], "wb");
    if (!f) return 2;
    fwrite(st, sizeof st[0], 2, f);
    save_cells(f, mu_SGS);
    fclose(f);
    fsl_layer2_detach();
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s run|filter|dynarg L H D ...\n", argv[0]);
        return 2;
    }
    LEN = (unsigned)atoi(argv[2]);
    HIG = (unsigned)atoi(argv[3]);
    DEP = (unsigned)atoi(argv[4]);
    if (!strcmp(argv[1], "run")) return run(argc, argv);
    if (!strcmp(argv[1], "filter")) return filter(argc, argv);
    if (!strcmp(argv[1], "dynarg")) return dynarg(argc, argv);
    return 2;
}

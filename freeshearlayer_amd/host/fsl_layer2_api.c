/*
 * fsl_layer2_api.c -- the reference's entry-point names over one bound
 * device context (ANSI C).  See include/fsl_layer2.h.
 *
 * The reference functions act on process globals (src/global.h); here the
 * globals' device copies live in the bound context and the three extra calls
 * of fsl_layer2.h move data at the points the reference's main needs it.
 */
#include "fsl_layer2.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static fsl_ctx *g_ctx = 0;
static fsl_case g_case;
static double ***g_U[5];
static double ***g_mu = 0;
static double *g_deltaT = 0;
static int g_status = FSL_OK;
static char g_msg[256];

static void note(int rc)
{
    if (rc != FSL_OK && g_status == FSL_OK) {
        g_status = rc;
        strncpy(g_msg, g_ctx ? fsl_last_error(g_ctx) : "no context attached", sizeof g_msg - 1);
        g_msg[sizeof g_msg - 1] = 0;
    }
}

static size_t gsize(const fsl_case *c) { return (size_t)(c->LEN + 2) * (c->HIG + 2) * (c->DEP + 2); }

/* real*** (rows allocated separately, src/helpers.c:58-80) <-> dense */
static void rows_to_dense(const fsl_case *c, double ***a, double *d)
{
    unsigned i, j;
    for (i = 0; i < c->LEN + 2; i++)
        for (j = 0; j < c->HIG + 2; j++)
            memcpy(d + ((size_t)i * (c->HIG + 2) + j) * (c->DEP + 2), a[i][j], (c->DEP + 2) * sizeof(double));
}

static void dense_to_rows(const fsl_case *c, const double *d, double ***a)
{
    unsigned i, j;
    for (i = 0; i < c->LEN + 2; i++)
        for (j = 0; j < c->HIG + 2; j++)
            memcpy(a[i][j], d + ((size_t)i * (c->HIG + 2) + j) * (c->DEP + 2), (c->DEP + 2) * sizeof(double));
}

int fsl_layer2_attach(const fsl_case *c, const fsl_config *cfg, double ***U1, double ***U2, double ***U3,
                      double ***U4, double ***U5, double ***mu_SGS, double *deltaT)
{
    double *buf;
    size_t G;
    int f, rc;
    fsl_layer2_detach();
    g_status = FSL_OK;
    g_msg[0] = 0;
    rc = fsl_create(c, cfg, &g_ctx);
    if (rc != FSL_OK) {
        g_ctx = 0;
        g_status = rc;
        strcpy(g_msg, "fsl_create failed");
        return rc;
    }
    g_case = *c;
    g_U[0] = U1; g_U[1] = U2; g_U[2] = U3; g_U[3] = U4; g_U[4] = U5;
    g_mu = mu_SGS;
    g_deltaT = deltaT;
    G = gsize(c);
    buf = (double *)malloc(5 * G * sizeof(double));
    if (!buf) return FSL_ERR_NOMEM;
    for (f = 0; f < 5; f++) rows_to_dense(c, g_U[f], buf + f * G);
    rc = fsl_upload_state(g_ctx, 0, buf);
    free(buf);
    if (rc == FSL_OK && deltaT) rc = fsl_set_dt(g_ctx, *deltaT);
    note(rc);
    return rc;
}

void fsl_layer2_detach(void)
{
    if (g_ctx) fsl_destroy(g_ctx);
    g_ctx = 0;
}

fsl_ctx *fsl_layer2_context(void) { return g_ctx; }

void fsl_layer2_set_block(unsigned k_min, unsigned k_max, double Ud, double Vd, double Wd)
{
    if (!g_ctx) { note(FSL_ERR_ARG); return; }
    note(fsl_set_block(g_ctx, k_min, k_max, Ud, Vd, Wd));
}

int fsl_layer2_sync_host(void)
{
    double *buf;
    size_t G;
    int f, rc;
    if (!g_ctx) return FSL_ERR_ARG;
    G = gsize(&g_case);
    buf = (double *)malloc(5 * G * sizeof(double));
    if (!buf) return FSL_ERR_NOMEM;
    rc = fsl_download_state(g_ctx, 0, buf);
    if (rc == FSL_OK)
        for (f = 0; f < 5; f++) dense_to_rows(&g_case, buf + f * G, g_U[f]);
    if (rc == FSL_OK && g_mu) {
        rc = fsl_download_mu_sgs(g_ctx, buf);
        if (rc == FSL_OK) dense_to_rows(&g_case, buf, g_mu);
    }
    free(buf);
    note(rc);
    return rc;
}

int fsl_layer2_status(void) { return g_status; }
const char *fsl_layer2_error(void) { return g_msg; }

#define NEED_CTX() do { if (!g_ctx) { note(FSL_ERR_ARG); return; } } while (0)

void BounCondInGhostCells(void) { NEED_CTX(); note(fsl_ghost_fill(g_ctx)); }
void Reconstruction(void) { NEED_CTX(); note(fsl_reconstruct(g_ctx)); }
void BounCondOnInterfaces(void) { NEED_CTX(); note(fsl_iface_bc(g_ctx)); }
void Fluxes(void) { NEED_CTX(); note(fsl_fluxes(g_ctx)); }
void Evolution(int numStages, int Stage) { NEED_CTX(); note(fsl_evolution(g_ctx, numStages, Stage)); }
void Evolution_threeStage_TVD_RK(int Stage) { NEED_CTX(); note(fsl_rk_update(g_ctx, 3, Stage)); }
void Evolution_twoStage_TVD_RK(int Stage) { NEED_CTX(); note(fsl_rk_update(g_ctx, 2, Stage)); }

void checkCoNum(void)
{
    double maxCo = 0.0;
    int rc;
    NEED_CTX();
    rc = fsl_check_co_num(g_ctx, &maxCo);
    /* src/helpers.c:44,48 */
    printf("Maximum Courant no. = %f\n", maxCo);
    printf("New time-step size:  %f\n", fsl_get_dt(g_ctx));
    if (g_deltaT) *g_deltaT = fsl_get_dt(g_ctx);
    note(rc);
}

double Random(double x) { return fsl_random(x); }
double minmod(double x, double y) { return fsl_minmod(x, y); }

double ***Array3D(int columns, int rows, int floors)
{
    double ***x;
    int i, j;
    x = (double ***)calloc((size_t)columns, sizeof(double **));
    if (!x) { puts("Cannot allocate memory"); exit(-1); }
    for (i = 0; i < columns; i++) {
        x[i] = (double **)calloc((size_t)rows, sizeof(double *));
        if (!x[i]) { puts("Cannot allocate memory"); exit(1); }
        for (j = 0; j < rows; j++) {
            x[i][j] = (double *)calloc((size_t)floors, sizeof(double));
            if (!x[i][j]) { puts("Cannot allocate memory"); exit(-1); }
        }
    }
    return x;
}

void free3D(double ***arr, int columns, int rows)
{
    int i, j;
    for (i = 0; i < columns; i++) {
        for (j = 0; j < rows; j++) free(arr[i][j]);
        free(arr[i]);
    }
    free(arr);
}

double ***filter_(double ***U, double h[3])
{
    const fsl_case *c = &g_case;
    size_t G;
    double *in, *out, ***r;
    int rc;
    if (!g_ctx) { note(FSL_ERR_ARG); return 0; }
    G = gsize(c);
    in = (double *)malloc(G * sizeof(double));
    out = (double *)malloc(G * sizeof(double));
    if (!in || !out) { free(in); free(out); note(FSL_ERR_NOMEM); return 0; }
    rows_to_dense(c, U, in);
    rc = fsl_filter(g_ctx, in, out, h);
    note(rc);
    r = Array3D((int)c->LEN + 2, (int)c->HIG + 2, (int)c->DEP + 2);
    if (rc == FSL_OK) dense_to_rows(c, out, r);
    free(in);
    free(out);
    return r;
}

void DynamicSmagorinsky(double ***rho, double ***ru, double ***rv, double ***rw, double ***mu_SGS)
{
    NEED_CTX();
    if (rho != g_U[0] || ru != g_U[1] || rv != g_U[2] || rw != g_U[3] || (g_mu && mu_SGS != g_mu)) {
        note(FSL_ERR_UNSUPPORTED);
        return;
    }
    note(fsl_dynamic_smagorinsky(g_ctx));
}

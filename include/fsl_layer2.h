/*
 * fsl_layer2.h -- the reference's own entry-point names over a bound device
 * context (libfsl_layer2.so, built on libfsl.so).
 *
 * A program written against the reference's src/ API (src/layer2.c) keeps its
 * main loop and links this library instead of the hot-path objects
 * (boundCondInGhostCells.o, reconstruction.o, bounCondOnInterfaces.o,
 * fluxes.o, turbulence.o, evolution.o, helpers.o).  It adds three calls:
 * fsl_layer2_attach() after Initialize(), fsl_layer2_set_block() after each
 * disturbance-block draw, fsl_layer2_sync_host() before Probes()/Output()/
 * Finalize().  The state then lives on the GPU.
 *
 * The functions keep the reference's void signatures; a failure is recorded
 * and returned by fsl_layer2_status() (the reference has no error returns).
 */
#ifndef FSL_LAYER2_H
#define FSL_LAYER2_H

#include "fsl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Creates a context for the case, uploads U1..U5 (the reference's real***
   arrays, real = double), remembers the arrays for fsl_layer2_sync_host(),
   and binds the context.  deltaT points at the caller's global deltaT, which
   checkCoNum() keeps up to date (may be NULL). */
int fsl_layer2_attach(const fsl_case *c, const fsl_config *cfg, double ***U1, double ***U2,
                      double ***U3, double ***U4, double ***U5, double ***mu_SGS, double *deltaT);
void fsl_layer2_detach(void);
fsl_ctx *fsl_layer2_context(void);
/* the caller's disturbance block (src/layer2.c:152-162 sets these globals) */
void fsl_layer2_set_block(unsigned k_min, unsigned k_max, double Ud, double Vd, double Wd);
/* copies the device step state (and mu_SGS) back into the attached arrays */
int fsl_layer2_sync_host(void);
/* first error since attach (FSL_OK if none) and its message */
int fsl_layer2_status(void);
const char *fsl_layer2_error(void);

/* --- reference names --- */
void BounCondInGhostCells(void);               /* src/boundCondInGhostCells.c:10 */
void Reconstruction(void);                     /* src/reconstruction.c:26 */
void BounCondOnInterfaces(void);               /* src/bounCondOnInterfaces.c:11 */
void Fluxes(void);                             /* src/fluxes.c:12 */
void Evolution(int numStages, int Stage);      /* src/evolution.c:15 */
void Evolution_threeStage_TVD_RK(int Stage);   /* src/evolution.c:29 */
void Evolution_twoStage_TVD_RK(int Stage);     /* src/evolution.c:164 */
void checkCoNum(void);                         /* src/helpers.c:23 */
double Random(double x);                       /* src/helpers.c:13 */
double minmod(double x, double y);             /* src/reconstruction.c:17 */
/* src/turbulence.c:17: host arrays in, newly allocated host array out */
double ***filter_(double ***U, double h[3]);
/* src/turbulence.c:103: the arguments must be the attached U1..U4, mu_SGS */
void DynamicSmagorinsky(double ***rho, double ***ru, double ***rv, double ***rw, double ***mu_SGS);
/* src/helpers.c:58, 85 (rows are zero-initialised here) */
double ***Array3D(int columns, int rows, int floors);
void free3D(double ***arr, int columns, int rows);

#ifdef __cplusplus
}
#endif
#endif

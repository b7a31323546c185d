/*
 * fsl_oracle.h -- CPU restatement of the freeShearLayer per-step update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; it is never linked into, loaded by or called from the product
 * library.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.
 *
 * It restates, in FP64 and in the reference's floating-point operation order,
 * the serial src/ path of nikola-m/freeShearLayer:
 *   src/boundCondInGhostCells.c:10-96      -> or_ghost_fill
 *   src/reconstruction.c:17-24,26-345      -> or_minmod, or_reconstruct
 *   src/bounCondOnInterfaces.c:11-110      -> or_iface_bc
 *   src/fluxes.c:12-573                    -> or_fluxes
 *   src/turbulence.c:17-100,103-507        -> or_filter, or_dynamic_smagorinsky
 *   src/evolution.c:15-245                 -> or_evolution
 *   src/helpers.c:13-17,23-52              -> or_random, or_check_co_num
 *   src/layer2.c:141-185                   -> or_step (stage driver, block draw)
 *   src/initialize.c:62-94                 -> or_params_derive
 *   src/output.c:88-136                    -> or_output_fields
 *   src/probes.c:17-40                     -> or_probe_values
 *
 * Every function here is pinned by tests/test_oracle_*.py: bit-for-bit against
 * the compiled reference (oracle/_ref, built from /root/reference/src by
 * oracle/Makefile) where that builds, against the committed golden fixtures
 * under tests/golden/ (written from the compiled reference), and against closed
 * forms.  Functions with no such pin say "parity unpinned" below.
 *
 * Layout (shared with the HIP library's dense upload/download format):
 *   cell array  [i][j][k], i in [0,L+1], j in [0,H+1], k in [0,D+1], k fastest
 *   x-face      [f][j][k], f in [0,L],   j in [0,H),   k in [0,D)
 *   y-face      [i][f][k], i in [0,L),   f in [0,H],   k in [0,D)
 *   z-face      [i][j][f], i in [0,L),   j in [0,H),   f in [0,D]
 * Face f in direction d lies between ghosted cells f and f+1 along d.
 */
#ifndef FSL_ORACLE_H
#define FSL_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_params {
    /* the positional case items of layer2.ini (src/initialize.c:28-49) */
    int L, H, D;                 /* LEN, HIG, DEP */
    double dx, dy, dz, dt;       /* deltaX, deltaY, deltaZ, deltaT */
    double mu_L, Pr_L, Pr_T, Cs; /* Cs holds Cs^2 (layer2.ini:13) */
    int BL_HIG;
    double Ua, Va, Wa;
    int Ns_min, Nst;
    int nStages;
    double maxCoNum;
    int dynamic;                 /* DynamicSmagorinskySGS (src/def.h:47) */
    int z_quirk;                 /* 1: src/reconstruction.c:328,340 as shipped */

    /* inflow state: ghost cells and face 0 (src/boundCondInGhostCells.c:17-31,
       src/bounCondOnInterfaces.c:17-62).  u_in[j-1] is the inflow velocity of
       cell row j; the reference uses 76.4 for j <= H/2 and 200 above. */
    double rho_in, p_in, p_out;
    const double *u_in;          /* H values */

    /* derived (src/initialize.c:67-94) */
    double rdx, rdy, rdz, r2dx, r2dy, r2dz, r4dx, r4dy, r4dz;
    double lambda_L, cp_Pr_T, DD, CsDD;

    /* time-varying scalars */
    double dtX, dtY, dtZ;        /* deltaT_X, deltaT_Y, deltaT_Z */
    unsigned k_min, k_max;       /* disturbance block (src/layer2.c:152-162) */
    double Ud, Vd, Wd;
    double X;                    /* LCG state */
    int counter;
} or_params;

typedef struct or_state {
    double *U[5];      /* U1..U5 (step state)              */
    double *Up[5];     /* U1p..U5p (stage state)           */
    double *xL[5], *xR[5];   /* xU*, U*x  : left/right state (fluxes overwrite xL) */
    double *yL[5], *yR[5];   /* yU*, U*y */
    double *zL[5], *zR[5];   /* zU*, U*z */
    double *mu_sgs;    /* mu_SGS, ghost layer stays 0 */
    int cur_is_p;      /* U_ -> U (0) or Up (1) */
} or_state;

/* sizes */
size_t or_ncell(const or_params *p);      /* (L+2)(H+2)(D+2) */
size_t or_nxface(const or_params *p);     /* (L+1) H D */
size_t or_nyface(const or_params *p);     /* L (H+1) D */
size_t or_nzface(const or_params *p);     /* L H (D+1) */

void or_params_default_ref(or_params *p, int L, int H, int D, const double *u_in_storage);
void or_params_derive(or_params *p);      /* src/initialize.c:67-94 */

int  or_state_alloc(const or_params *p, or_state *s);   /* zero-filled */
void or_state_free(or_state *s);

void or_ic_reference(const or_params *p, or_state *s);   /* src/initialize.c:159-187 */

double or_random(double x);                              /* src/helpers.c:13-17 */
double or_minmod(double x, double y);                    /* src/reconstruction.c:17-24 */
void   or_block_draw(or_params *p);                      /* src/layer2.c:152-162 */
double or_check_co_num(or_params *p, const or_state *s); /* src/helpers.c:23-52, returns maxCo */

void or_ghost_fill(const or_params *p, double *const U[5]);
void or_reconstruct(const or_params *p, or_state *s);
void or_iface_bc(const or_params *p, or_state *s);
void or_fluxes(const or_params *p, or_state *s);         /* calls or_dynamic_smagorinsky when p->dynamic */
void or_evolution(const or_params *p, or_state *s, int numStages, int stage);

/* filter_ (src/turbulence.c:17-100): out must be zero-filled on entry */
void or_filter(const or_params *p, const double *in, double *out, const double h[3]);
void or_dynamic_smagorinsky(const or_params *p, double *const U[5], double *mu_sgs);
/* the same model's Germano terms L:M, M:M and rho DD |S| per interior cell
   (G arrays; test reference of the averaged coefficients) */
void or_dynamic_terms(const or_params *p, double *const U[5], double *lm, double *mm, double *fac);

/* One time step in the order of src/layer2.c:141-185 (probes/output left out). */
void or_step(or_params *p, or_state *s);

/* The nine computed columns of Output()'s Tecplot rows (src/output.c:88-136:
   rho, u, v, w, p, T, vorticity magnitude, Q, rho CsDD |S| / mu_L) of the
   step state for every interior cell, i fastest, then j, then k
   (9 L H D values). */
void or_output_fields(const or_params *p, double *const U[5], double *out);
/* Probes() columns (src/probes.c:17-40): U1 and U2 along j = 1..H at
   i = L/2, 3L/4, L on k = D/2 + 1 (6 H values, in the file's order). */
void or_probe_values(const or_params *p, double *const U[5], double *out);
/* Fault-injection hook for tests only: when nonzero, or_fluxes drops the heat
   flux term.  Used to show the parity tests catch a dropped term. */
extern int or_fault_drop_heat_flux;

/* ---- handle surface for ctypes (tests, bench cpu_baseline) ---- */
typedef struct or_sim or_sim;
or_sim *or_sim_create(int L, int H, int D, int dynamic);
void    or_sim_destroy(or_sim *s);
int     or_sim_set(or_sim *s, const char *key, double value);   /* 0 ok, -1 unknown key */
double  or_sim_get(or_sim *s, const char *key);
void    or_sim_derive(or_sim *s);
double *or_sim_u_in(or_sim *s);                                 /* H inflow velocities */
/* cell arrays: 0..4 U1..U5, 5..9 U1p..U5p, 10 mu_SGS */
double *or_sim_cell(or_sim *s, int id);
/* face arrays: 0..4 xU*, 5..9 U*x, 10..14 yU*, 15..19 U*y, 20..24 zU*, 25..29 U*z */
double *or_sim_face(or_sim *s, int id);
void    or_sim_ghost(or_sim *s);
void    or_sim_recon(or_sim *s);
void    or_sim_iface(or_sim *s);
void    or_sim_fluxes(or_sim *s);
void    or_sim_evolution(or_sim *s, int numStages, int stage);
double  or_sim_check_co_num(or_sim *s);
void    or_sim_block_draw(or_sim *s);
void    or_sim_dynamic(or_sim *s);
void    or_sim_dynamic_terms(or_sim *s, double *lm, double *mm, double *fac);
void    or_sim_filter(or_sim *s, const double *in, double *out);
void    or_sim_step(or_sim *s);
void    or_sim_ic_reference(or_sim *s);
int     or_sim_cur_is_p(or_sim *s);
void    or_sim_set_cur(or_sim *s, int is_p);
void    or_sim_output_fields(or_sim *s, double *out);
void    or_sim_probe_values(or_sim *s, double *out);

#ifdef __cplusplus
}
#endif
#endif

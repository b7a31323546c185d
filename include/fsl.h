/*
 * fsl.h -- C-ABI of the MI355X-native freeShearLayer per-step update.
 *
 * Plain C types only (pointers, sizes, doubles).  Two layers:
 *
 *   1. Case input and host-side helpers (ANSI C, libfsl.so):
 *      .ini / .bin readers, the inflow LCG and disturbance block, initial
 *      conditions, backup.dat I/O.
 *   2. Device context (HIP on gfx950, libfsl.so): device-resident state and
 *      the per-stage operations of the reference's solver step.  Each
 *      operation names the reference function it replaces (file:line in
 *      nikola-m/freeShearLayer).
 *
 * The reference-named drop-in entry points (Reconstruction(), Fluxes(),
 * Evolution(), ...) operating on a bound context are declared in
 * fsl_layer2.h and live in libfsl_layer2.so.
 *
 * Conventions
 *   - Every fsl_* function that can fail returns an int status (FSL_OK = 0,
 *     negative on error); fsl_last_error() gives a message.
 *   - Host arrays use the dense layouts
 *       cell  [i][j][k]  (LEN+2)(HIG+2)(DEP+2), k fastest, one ghost layer
 *       x-face [f][j][k] (LEN+1) HIG DEP
 *       y-face [i][f][k] LEN (HIG+1) DEP
 *       z-face [i][j][f] LEN HIG (DEP+1)
 *     A "state" is the five conserved fields rho, rho u, rho v, rho w, E
 *     stored one after the other (5 cell arrays).
 *   - A context is single-threaded: calls on one context must not overlap.
 *     Work is queued on the context's own HIP stream; functions that return
 *     host data synchronise that stream.
 */
#ifndef FSL_H
#define FSL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSL_OK            0
#define FSL_ERR_ARG      -1
#define FSL_ERR_HIP      -2
#define FSL_ERR_NOMEM    -3
#define FSL_ERR_NONFINITE -4   /* CFL reduction gave 0, inf or NaN */
#define FSL_ERR_IO       -5
#define FSL_ERR_UNSUPPORTED -6

/* ------------------------------------------------------------------ */
/* Case input                                                          */
/* ------------------------------------------------------------------ */

/* The 22 positional items of layer2.ini, in file order
   (src/initialize.c:18-58; the same order as the 22 float32 of
   mpi_layer2.bin, src-par/initialize.c:53-77). */
typedef struct fsl_case {
    unsigned LEN, HIG, DEP;
    double deltaX, deltaY, deltaZ, deltaT;
    unsigned Answer, numstep;
    double mu_L, Pr_L, Pr_T;
    double Cs;                 /* holds Cs^2 (layer2.ini:13) */
    unsigned BL_HIG;
    double Ua, Va, Wa;
    int Ns_min, Nst;
    unsigned f_step;
    int nStages;
    double maxCoNum;
} fsl_case;

/* Replaces the parse of src/initialize.c:18-60: 22 lines read with
   fgets(buf, 79, f), atoi/atof of each line.  FSL_ERR_IO if a line is missing. */
int fsl_case_read_ini(const char *path, fsl_case *out);
/* src-par/initialize.c:26-85: 22 native float32 values, integers converted
   from float. */
int fsl_case_read_bin(const char *path, fsl_case *out);
/* src-par/ini2bin/ini2bin.c:11-81: .ini -> 22 float32 (atof of each line). */
int fsl_case_ini_to_bin(const char *ini_path, const char *bin_path);
/* Checks the case for sizes the reference would index out of bounds with
   (BL_HIG > HIG - HIG/2, src/bounCondOnInterfaces.c:31) and for zeros. */
int fsl_case_validate(const fsl_case *c, char *msg, size_t msglen);

/* ------------------------------------------------------------------ */
/* Host helpers                                                        */
/* ------------------------------------------------------------------ */

/* src/helpers.c:13-17: x <- frac(11 x + pi). */
double fsl_random(double x);
/* src/reconstruction.c:17-24 on the host (the device copy is inlined). */
double fsl_minmod(double x, double y);

/* Disturbance block of src/layer2.c:152-162, advanced once per step. */
typedef struct fsl_block {
    double X;          /* LCG state (seed: the number on the reference's stdin) */
    int counter;       /* steps left before the next draw */
    unsigned k_min, k_max;
    double Ud, Vd, Wd;
} fsl_block;
void fsl_block_init(fsl_block *b, double seed);
/* returns 1 when a new block was drawn */
int fsl_block_advance(fsl_block *b, const fsl_case *c);

/* Initial conditions into a host state (5 cell arrays, ghosts zeroed). */
/* src/initialize.c:159-187: two streams, 76.4 m/s for j <= HIG/2, 200 above. */
void fsl_ic_reference(const fsl_case *c, double *U);

/* Spatially developing shear layer of the synthetic benchmark case:
   U(y) = Um + dU/2 tanh(2 y / delta_w0), y measured from the j = HIG/2 face,
   uniform p0 and rho0, and seeded uniform perturbations of amplitude
   pert * dU on u, v, w where |y| < delta_w0 (splitmix64 of seed ^ cell index).
   E is rebuilt from p0.  u_in (HIG values, may be NULL) receives U(y_j),
   the matching inflow profile. */
typedef struct fsl_shear_case {
    double mach_c;     /* convective Mach number Mc */
    double ratio;      /* U2 / U1 */
    double delta_w0;   /* vorticity thickness, m */
    double reynolds;   /* Re on delta_w0 (sets mu_L = rho0 dU delta_w0 / Re) */
    double pert;       /* perturbation amplitude as a fraction of dU */
    double rho0, p0;
    unsigned long long seed;
} fsl_shear_case;
void fsl_shear_defaults(fsl_shear_case *s);
/* fills U and u_in; returns the mu_L the case implies through Re */
double fsl_ic_shear_layer(const fsl_case *c, const fsl_shear_case *s, double *U, double *u_in);

/* backup.dat of src/finalize.c:14-34: for i, j: DEP values of each of U1..U5,
   then deltaT, then step (uint32). */
int fsl_backup_write(const char *path, const fsl_case *c, const double *U, double deltaT,
                     unsigned step, int real_bytes /* 8: real = double, 4: float as shipped */);
/* Reader of src/initialize.c:133-158.  The reference reads step from the
   bytes right after the fields (where the writer put deltaT); ref_layout = 1
   reproduces that, 0 reads deltaT then step as written. */
int fsl_backup_read(const char *path, const fsl_case *c, double *U, double *deltaT,
                    unsigned *step, int ref_layout, int real_bytes);

/* ------------------------------------------------------------------ */
/* Device context                                                      */
/* ------------------------------------------------------------------ */

typedef struct fsl_ctx fsl_ctx;

#define FSL_SGS_CONSTANT 0   /* src/fluxes.c:195-203, DynamicSmagorinskySGS = 0 */
#define FSL_SGS_DYNAMIC  1   /* src/turbulence.c:103-507, as shipped */
#define FSL_SGS_LEARNED  2   /* per-cell MLP for Cs^2, dynamic model as judge */

/* Averaging of the dynamic coefficient (fsl_set_sgs_average).  The
   reference's model is pointwise (NONE, the default and the parity mode).
   Z: Cd(i, j) = -1/2 sum_k L:M / (sum_k M:M + 1e-30) over each z line;
   XZ: Cd(j) from the sums over each x-z plane (the "plane averaging" of the
   Germano-Lilly procedure); both clipped to [0, 0.15], mu_SGS = rho Cd
   Delta^2 |S|.  Applies to the dynamic model wherever it runs (the dynamic
   path and the learned closure's judge and fallback). */
#define FSL_SGS_AVG_NONE 0
#define FSL_SGS_AVG_Z    1
#define FSL_SGS_AVG_XZ   2

#define FSL_PREC_FP64    0   /* parity path: FP64, no contraction, reference op order */
#define FSL_PREC_MIXED   1   /* FP32 stencils and SGS fluxes, FP64 state / RK / p recovery /
                                CFL / dynamic model */
#define FSL_PREC_FP64_FAST 2 /* FP64 with the fused stage kernel built with FMA contraction
                                and reciprocal quotients (needs fused = 1); within 1e-12
                                relative of the parity path after one step, not bit-identical;
                                every other kernel is the parity path's */

#define FSL_CFL_CONVECTIVE 0 /* src/helpers.c:23-52: max (rho u / rho) dt/dx, rescale */
#define FSL_CFL_ACOUSTIC   1 /* dt = cfl / max[(|u|+c)/dx + (|v|+c)/dy + (|w|+c)/dz] */

typedef struct fsl_config {
    int sgs_model;        /* FSL_SGS_* */
    int precision;        /* FSL_PREC_* */
    int cfl_mode;         /* FSL_CFL_* */
    double cfl;           /* acoustic CFL number (convective mode uses maxCoNum) */
    int z_quirk;          /* 1: z reconstruction as shipped (src/reconstruction.c:328,340) */
    int fused;            /* 1: fused stage kernels (3 stages only); 0: one kernel per entry point */
    double rho_in, p_in;  /* inflow state (reference: 1.0, 1e5) */
    double p_out;         /* outflow pressure (reference: 1e5) */
    int device;           /* HIP device ordinal */
} fsl_config;

/* Reference defaults: dynamic SGS, FP64, convective CFL, z quirk, unfused,
   inflow 1.0 / 1e5 / 1e5, device 0. */
void fsl_config_defaults(fsl_config *cfg);

/* Allocates device state for the case (src/initialize.c:96-125 allocates the
   reference's 51 host arrays).  The inflow profile defaults to the reference
   step (76.4 / 200 m/s). */
int fsl_create(const fsl_case *c, const fsl_config *cfg, fsl_ctx **out);
void fsl_destroy(fsl_ctx *ctx);
const char *fsl_last_error(const fsl_ctx *ctx);
/* sizes: cells (LEN+2)(HIG+2)(DEP+2), faces per direction */
size_t fsl_ncell(const fsl_ctx *ctx);
size_t fsl_nface(const fsl_ctx *ctx, int dir);
/* bytes of device memory the context holds */
size_t fsl_device_bytes(const fsl_ctx *ctx);

/* inflow velocity per cell row j = 1..HIG (HIG values) */
int fsl_set_inflow(fsl_ctx *ctx, const double *u_in);

/* State transfer.  which: 0 = U (step state), 1 = Up (stage state).
   Layout: 5 dense cell arrays.  Upload after Initialize / restart; download
   before Output / Probes / Finalize. */
int fsl_upload_state(fsl_ctx *ctx, int which, const double *U);
int fsl_download_state(fsl_ctx *ctx, int which, double *U);
/* Face arrays of the unfused path: id 0..4 xU*, 5..9 U*x, 10..14 yU*,
   15..19 U*y, 20..24 zU*, 25..29 U*z (after Fluxes, the xU*, yU*, zU* arrays hold the
   fluxes, as src/def.h:6-22 aliases them). */
int fsl_download_face(fsl_ctx *ctx, int id, double *out);
int fsl_upload_face(fsl_ctx *ctx, int id, const double *in);
/* mu_SGS cell array (ghost layer 0) */
int fsl_download_mu_sgs(fsl_ctx *ctx, double *out);
int fsl_upload_mu_sgs(fsl_ctx *ctx, const double *in);

/* Scalars.  fsl_set_dt sets deltaT and deltaT_X/Y/Z = deltaT / delta. */
int fsl_set_dt(fsl_ctx *ctx, double deltaT);
double fsl_get_dt(const fsl_ctx *ctx);
int fsl_set_block(fsl_ctx *ctx, unsigned k_min, unsigned k_max, double Ud, double Vd, double Wd);
/* stage-input selector of the reference (U_ = U or U_ = Up) */
int fsl_get_cur(const fsl_ctx *ctx);
int fsl_set_cur(fsl_ctx *ctx, int is_p);

/* --- per-entry-point operations (unfused path) --- */
/* src/boundCondInGhostCells.c:10-96 on the stage input U_ */
int fsl_ghost_fill(fsl_ctx *ctx);
/* src/reconstruction.c:26-345: 30 face arrays from U_ */
int fsl_reconstruct(fsl_ctx *ctx);
/* src/bounCondOnInterfaces.c:11-110 */
int fsl_iface_bc(fsl_ctx *ctx);
/* src/fluxes.c:12-573 (calls the dynamic model first when configured) */
int fsl_fluxes(fsl_ctx *ctx);
/* src/evolution.c:15-27 (2 stages falls through into the 3-stage update) */
int fsl_evolution(fsl_ctx *ctx, int numStages, int stage);
/* One scheme's update only: scheme 3 = src/evolution.c:29-161,
   scheme 2 = src/evolution.c:164-245 */
int fsl_rk_update(fsl_ctx *ctx, int scheme, int stage);
/* src/helpers.c:23-52: device max-reduction, then deltaT *= maxCoNum / maxCo
   (convective) or deltaT = cfl / max(...) (acoustic).  maxCo may be NULL. */
int fsl_check_co_num(fsl_ctx *ctx, double *maxCo);
/* src/turbulence.c:103-507 on the step state U -> mu_SGS: the reference's
   FP64 operations in every context (mixed-precision contexts with the
   environment variable FSL_SGS_FP32=1: the same model with FP32 filters,
   contraction and scratch) */
int fsl_dynamic_smagorinsky(fsl_ctx *ctx);
/* Averaging of the dynamic coefficient, FSL_SGS_AVG_* (no reference
   counterpart: the reference's model is pointwise; default NONE).  The
   averaged modes keep 2 G + 2 LEN HIG + 2 HIG doubles of terms and sums. */
int fsl_set_sgs_average(fsl_ctx *ctx, int mode);
int fsl_get_sgs_average(const fsl_ctx *ctx);
/* src/turbulence.c:17-100 with kernel h (the model uses {1/3,1/3,1/3}):
   host cell array in -> host cell array out (the returned array of the
   reference) */
int fsl_filter(fsl_ctx *ctx, const double *in, double *out, const double h[3]);

/* --- fused path --- */
/* One RK stage (ghost fill, reconstruction, interface BC, fluxes, update)
   with fused kernels; stage in 1..3. */
int fsl_stage(fsl_ctx *ctx, int stage);

/* --- driver (ANSI C over the operations above) --- */
/* One time step in the order of src/layer2.c:141-185: CFL rescale, block
   draw, nStages stages.  The block state lives in the context. */
int fsl_step(fsl_ctx *ctx);
int fsl_run(fsl_ctx *ctx, int nsteps);
int fsl_set_lcg(fsl_ctx *ctx, double X, int counter);
int fsl_get_lcg(const fsl_ctx *ctx, double *X, int *counter, unsigned *k_min, unsigned *k_max,
                double *Ud, double *Vd, double *Wd);
unsigned fsl_get_step(const fsl_ctx *ctx);
int fsl_set_step(fsl_ctx *ctx, unsigned step);
int fsl_get_block_state(const fsl_ctx *ctx, fsl_block *b);
int fsl_set_block_state(fsl_ctx *ctx, const fsl_block *b);
int fsl_get_case(const fsl_ctx *ctx, fsl_case *out);
int fsl_get_config(const fsl_ctx *ctx, fsl_config *out);
int fsl_synchronize(fsl_ctx *ctx);

/* --- diagnostics (device FP64 reductions) --- */
/* out[0..4]: sums of rho, rho u, rho v, rho w, E over the interior */
int fsl_conservation_sums(fsl_ctx *ctx, double out[5]);
/* The same sums without a context and without allocating: no counterpart in
   the reference; for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   U: device pointer to a state in the context's layout (5 dense cell arrays
   (LEN+2)(HIG+2)(DEP+2), one after the other); sums: device pointer to 5
   doubles, written in stream order; stream: a hipStream_t (NULL: the null
   stream).  The work is queued, not waited for.  The per-block partials live
   in a library-owned device array of 1024 x 5 doubles (40 KiB per device,
   held until the process ends); the partial count, min(ceil(N / 256), 1024),
   always fits.  Calls that overlap on one device share that array and race:
   serialise them (one stream, or events).  Same kernels and summation order
   as fsl_conservation_sums, so the results are bitwise equal. */
int fsl_conservation_sums_noalloc(const double *U, unsigned LEN, unsigned HIG, unsigned DEP, double *sums,
                                  void *stream);
/* Plane profiles of the step state: z sums of rho, rho u, rho v, rho w,
   (rho u)^2/rho, (rho v)^2/rho, (rho w)^2/rho for every interior (i, j);
   out[q LEN HIG + (i-1) HIG + (j-1)], q = 0..6 (7 LEN HIG doubles). */
int fsl_plane_profiles(fsl_ctx *ctx, double *out);

/* Shear-layer statistics from those profiles (host).  With the Favre mean
   u~ = sum(rho u) / sum(rho) and rho_bar = sum(rho) / DEP of each (i, j):
     delta_theta(x_i) = 1 / (rho0 dU^2) sum_j rho_bar (U1 - u~)(u~ - U2) deltaY
     delta_w(x_i)     = dU / max_j |u~(j+1) - u~(j-1)| / (2 deltaY)
     tke              = sum_{i,j} 1/2 sum_c [(rho u_c)^2/rho sums - (sum rho u_c)^2 / sum rho] * cell volume
   with dU = U1 - U2 (U1 the faster stream).  Averages of delta_theta and
   delta_w are over the planes i > LEN / 2.  dtheta_x (LEN values) and dw_x
   may be NULL. */
typedef struct fsl_shear_stats {
    double delta_theta;      /* mean over i > LEN/2 */
    double delta_theta_out;  /* at i = LEN */
    double delta_w;          /* mean over i > LEN/2 */
    double tke;              /* resolved (Favre) TKE integrated over the domain, J */
} fsl_shear_stats;
int fsl_shear_statistics(const fsl_case *c, const double *prof, double U1, double U2, double rho0,
                         double *dtheta_x, double *dw_x, fsl_shear_stats *out);

/* --- output files of the reference (src/output.c, src/probes.c,
       favre-average-src/layer2_p.c), from the device step state --- */
/* The nine computed columns of the Tecplot frame of src/output.c:88-136 for
   the interior planes k0 .. k0 + nk - 1 (1-based): rho, u, v, w, p, T,
   vorticity magnitude, Q = (|omega|^2 - |S|^2)/2 and mu_T / mu_L (the
   constant-coefficient rho CsDD |S| / mu_L, as src/output.c:134 writes in
   every SGS mode), FP64 in the reference's operation order; out holds
   9 values per cell, cells in the frame's order (i fastest, then j, k),
   nk LEN HIG cells. */
int fsl_output_fields(fsl_ctx *ctx, int k0, int nk, double *out);
/* Replaces Output() (src/output.c:16-143): writes the frame of the step
   state to path (NULL: "<step>.plt" in the working directory, as the
   reference names it), formatted with the reference's "%g" conversions. */
int fsl_write_plt(fsl_ctx *ctx, const char *path);
/* The same frame from host fields (9 values per interior cell in the
   frame's order, LEN HIG DEP cells, as fsl_output_fields gives them). */
int fsl_plt_write_fields(const char *path, const fsl_case *c, const double *fields);
/* Probe columns of src/probes.c:20-40: rho and rho u along j = 1..HIG at
   i = LEN/2, 3 LEN/4 and LEN on the plane k = DEP/2 + 1, in that order
   (6 HIG values: rho, rho u of station 1, then 2, then 3). */
int fsl_probe_values(fsl_ctx *ctx, double *out);
/* Replaces Probes() (src/probes.c:13-43): appends the step (uint32) and the
   six columns as `real` (real_bytes 8: double, 4: float as shipped) to path
   (NULL: "probes.dat"). */
int fsl_write_probes(fsl_ctx *ctx, const char *path, int real_bytes);
int fsl_probes_append(const char *path, unsigned step, unsigned HIG, const double *vals, int real_bytes);
/* favre-average-src/layer2_p.c:27-113: from a probes file, the sums of rho
   and rho u over the records with step_start <= step <= step_end (float
   arithmetic for the 4-byte layout, as the post-processor does), then the
   Favre mean u~ = sum(rho u) / sum(rho) of each station and row:
   ru[s HIG + j], s = 0..2.  Reading stops at the first record past step_end
   or at the end of the file; nrec (may be NULL) receives the records used. */
int fsl_favre_average(const char *path, unsigned HIG, unsigned step_start, unsigned step_end, int real_bytes,
                      double *ru, unsigned *nrec);
/* favre-average-src/layer2_p.c:115-147: probe1.dat .. probe3.dat in dir
   (NULL or "": the working directory). */
int fsl_favre_write_profiles(const char *dir, unsigned HIG, const double *ru);

/* --- learned SGS closure (sgs_model = FSL_SGS_LEARNED) --- */
/* Network 8 -> 64 -> 64 -> 1 with tanh or ReLU hidden layers
   (fsl_mlp_set_activation); BF16 operands on the matrix cores with FP32
   accumulation for training and inference.  Output y gives Cd = clip(0.15 y, 0,
   0.15) and mu_SGS = rho Cd Delta^2 |S| (src/turbulence.c:436-445 with the
   network's Cd in place of the Germano-Lilly one).  Features per interior
   cell, from central differences of u = (rho u)/rho (FP32):
     0 (|W|^2 - |S|^2) / (|W|^2 + |S|^2)    4 Delta |lap u| / |S| (second differences)
     1 tr(S^3) / |S|_F^3                     5 omega.S.omega / (|omega|^2 |S|)
     2 tr(S W^2) / (|S|_F |W|_F^2)          6 |S| Delta / dU
     3 tr(S^2 W^2) / (|S|_F^2 |W|_F^2)      7 y / delta_w(x)
   with |S| = sqrt(2 S:S), |S|_F^2 = S:S, Delta = (dx dy dz)^(1/3), dU the
   inflow velocity difference
   and delta_w(x) the vorticity thickness of the x plane, refreshed at every
   judge.  Parameters: FP32 vector W1[8][64], b1[64], W2[64][64], b2[64],
   w3[64], b3 (FSL_MLP_NPARAM values); norm[16]: feature means, then
   reciprocal standard deviations (the network sees (f - mean) * rstd). */
#define FSL_MLP_NIN 8
#define FSL_MLP_NHID 64
#define FSL_MLP_NPARAM 4801
/* host: Glorot-uniform weights (splitmix64 of seed), zero biases, identity norm */
void fsl_mlp_init_params(unsigned long long seed, float *params, float *norm);
/* Hidden-layer activation: tanh (the default) or ReLU h = max(z, 0).  The
   inference of a ReLU network takes one packed 16-bit maximum per two hidden
   units where tanh takes v_exp_f32, v_rcp_f32 and an add per unit. */
#define FSL_MLP_TANH 0
#define FSL_MLP_RELU 1
/* parameter file: "FSLMLP01", int32 nin, nhid, nparam, float norm[16], float
   params[nparam] (tanh); "FSLMLP02": int32 nin, nhid, nparam, activation, then
   the same.  fsl_mlp_file_write writes FSLMLP01; fsl_mlp_file_read reads
   either and refuses (FSL_ERR_UNSUPPORTED) a file whose activation is not
   tanh; the _act forms carry the activation. */
int fsl_mlp_file_write(const char *path, const float *params, const float *norm);
int fsl_mlp_file_read(const char *path, float *params, float *norm);
int fsl_mlp_file_write_act(const char *path, const float *params, const float *norm, int activation);
int fsl_mlp_file_read_act(const char *path, float *params, float *norm, int *activation);
int fsl_mlp_set_params(fsl_ctx *ctx, const float *params, const float *norm);
/* the context network's activation (default FSL_MLP_TANH); training and
   inference both use it */
int fsl_mlp_set_activation(fsl_ctx *ctx, int activation);
int fsl_mlp_get_activation(const fsl_ctx *ctx);
int fsl_mlp_get_params(fsl_ctx *ctx, float *params, float *norm);
/* network output for n host rows of 8 raw features (normalised with norm) */
int fsl_mlp_forward(fsl_ctx *ctx, const float *X, long long n, float *y);
/* raw features of interior cells offset, offset + stride, ... (n of them;
   interior order k fastest, then j, then i) of the step state U */
int fsl_mlp_features(fsl_ctx *ctx, long long stride, long long offset, long long n, float *X);
/* delta_w(x) from U; read (LEN+2 floats) or set it */
int fsl_mlp_update_thickness(fsl_ctx *ctx);
int fsl_mlp_get_thickness(fsl_ctx *ctx, float *dw);
int fsl_mlp_set_thickness(fsl_ctx *ctx, const float *dw);
/* learned mu_SGS of the step state U (ghost layer filled by the caller) */
int fsl_learned_smagorinsky(fsl_ctx *ctx);
/* Training-sample store on the device.  fsl_mlp_collect fills the ghost
   layer of U, runs the dynamic model and appends the features and targets
   Cd / 0.15 of interior cells offset, offset + stride, ... up to the
   reserved capacity (mu_SGS is left holding the dynamic model's values). */
int fsl_mlp_reserve(fsl_ctx *ctx, long long cap);
int fsl_mlp_collect(fsl_ctx *ctx, long long stride, long long offset, long long *n_total);
long long fsl_mlp_sample_count(const fsl_ctx *ctx);
int fsl_mlp_clear_samples(fsl_ctx *ctx);
int fsl_mlp_samples(fsl_ctx *ctx, long long first, long long n, float *X, float *t);
int fsl_mlp_load_samples(fsl_ctx *ctx, const float *X, const float *t, long long n);
/* feature means and reciprocal standard deviations of the store -> norm[16] */
int fsl_mlp_fit_norm(fsl_ctx *ctx, float *norm);
/* Sample weights of the loss: (|S| Delta / dU)^power (feature 6) scaled to
   mean 1 over the store; power 3 weights samples by their share of the SGS
   dissipation rho Cd Delta^2 |S|^3.  Default power 0 (unweighted). */
int fsl_mlp_set_loss_weight(fsl_ctx *ctx, float power);
/* Adam (0.9, 0.999, 1e-8) on the weighted mean squared error of y against the
   targets; epochs over a fresh permutation each, batch a multiple of 256;
   epoch_loss (epochs values, may be NULL) receives each epoch's mean loss. */
int fsl_mlp_train(fsl_ctx *ctx, int epochs, long long batch, float lr, unsigned long long seed,
                  float *epoch_loss);
/* gradient of the weighted mean squared error over samples 0 .. batch-1
   (batch a multiple of 256) at the current parameters, without an update;
   loss receives the sum of weighted squared errors */
int fsl_mlp_gradient(fsl_ctx *ctx, long long batch, float *grad, float *loss);
/* Judge: every interval steps (steps 1, 1 + interval, ...) both models run
   on the step state and three errors of the network's mu_SGS against the
   dynamic model's are formed, with e = mu_SGS |S|^2 the SGS dissipation
   density and D its sum over the interior: the level error |D_l - D_d| /
   D_d; the profile error || s P_l - P_d ||_2 / || P_d ||_2 of the y
   profiles P(j) (x-z sums of e in each row) and the local error
   || s e_l - e_d ||_2 / || e_d ||_2 over the cells, both with s = D_d / D_l
   (the distribution once the level is matched).  When the level error
   exceeds tol, or the profile or local error its bound
   (fsl_learned_set_tolerances; default 1e30: reported, not enforced), the
   step runs on the dynamic model, and so does every step until a later
   judge passes.  With adapt 1, each judge also sets the network's output
   scale (Cd = clip(0.15 scale y, 0, 0.15)) to scale D_d / D_l, clamped to
   [0.1, 10], for the following steps: the network gives the spatial
   distribution of Cd, the dynamic procedure its global level.  With adapt 2
   the level is set per y row, with adapt 3 per x band and y row (the bands
   of the judge's sums: NB = min(16, LEN), band b the planes b LEN / NB + 1
   .. (b + 1) LEN / NB): cell (i, j)'s scale is scale f(i, j), and each judge
   multiplies the factor of band b and row j by (P_d(b, j) + eps) /
   (P_l(b, j) + eps) / (D_d / D_l), eps = 1 % of the mean dissipation of a
   band's row, clamped to [0.1, 10] (rows outside the layer, of negligible
   dissipation, keep their factor); f(i, j) is linear in x between the band
   centres (adapt 2: one band).  Defaults: interval 50, tol 0.10, scale 1,
   adapt 2; interval 0 never judges; adapt 0 keeps the scale. */
int fsl_learned_configure(fsl_ctx *ctx, int interval, double tol);
int fsl_learned_set_tolerances(fsl_ctx *ctx, double profile_tol, double local_tol);
/* sets the scale and the adapt mode (0 .. 3); adapt 2 and 3 reset the level factors to 1 */
int fsl_learned_set_scale(fsl_ctx *ctx, double scale, int adapt);
/* level factors f(i, j) of the interior, out[(i - 1) HIG + j - 1] for
   q < nmax (all 1 below adapt 2); returns LEN HIG */
long long fsl_learned_level_factors(const fsl_ctx *ctx, long long nmax, float *out);
typedef struct fsl_learned_status {
    int interval;
    double tol;
    int in_fallback;
    double last_error;       /* -1 before the first judge */
    long long judged, fallbacks, learned_steps, dynamic_steps;
    double scale;            /* output scale in use */
    int adapt;
    double profile_tol, local_tol;
    double last_profile_error;   /* -1 before the first judge */
    double last_local_error;
} fsl_learned_status;
int fsl_learned_get_status(const fsl_ctx *ctx, fsl_learned_status *s);
/* judge history: error and the scale in use at each judge; returns the count */
int fsl_learned_history(const fsl_ctx *ctx, int nmax, double *err, double *scale);
/* the same with the profile and local errors (any pointer may be NULL) */
int fsl_learned_history_errors(const fsl_ctx *ctx, int nmax, double *err, double *profile_err, double *local_err,
                               double *scale);
/* one judge evaluation on U now (no state change beyond mu_SGS = dynamic) */
int fsl_learned_judge(fsl_ctx *ctx, double *err, double *d_dynamic, double *d_learned);
int fsl_learned_judge_errors(fsl_ctx *ctx, double *err, double *profile_err, double *local_err,
                             double *d_dynamic, double *d_learned);
/* one v_mfma_f32_16x16x16_bf16 product D = A B of row-major 16x16 FP32
   matrices rounded to BF16 (operand layout check) */
int fsl_mfma_probe(fsl_ctx *ctx, const float *A, const float *B, float *D);

/* --- measurement --- */
/* When on, every kernel launch is bracketed by HIP events on the context's
   stream; fsl_profile_read returns, per kernel id, the accumulated milliseconds
   and launch count.  Names by fsl_kernel_name(id). */
int fsl_profile_enable(fsl_ctx *ctx, int on);
int fsl_profile_reset(fsl_ctx *ctx);
int fsl_profile_read(fsl_ctx *ctx, int nmax, double *ms, long long *count);
const char *fsl_kernel_name(int id);
int fsl_kernel_count(void);
/* device copy of n doubles (calibration of HBM counters against a plain copy) */
int fsl_copy_kernel(fsl_ctx *ctx, size_t n, int reps);

#ifdef __cplusplus
}
#endif
#endif
